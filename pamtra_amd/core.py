"""The pyPamtra user API (python/pyPamtra/core.py:74-1568), passive path.

Same names and dictionaries as the reference: ``pyPamtra()`` with ``nmlSet``,
``set``, ``df`` (descriptor table), ``p`` (profile), ``r`` (results),
``createProfile(**kwargs)`` and ``runPamtra(freqs)``.  ``runPamtra`` hands the
dictionaries to ``libWrapper.PamtraFortranWrapper``, which runs the HIP
library instead of the f2py Fortran module.
"""

from __future__ import annotations

import time
import warnings
from copy import deepcopy

import numpy as np

from .descriptorFile import pamDescriptorFile
from .libWrapper import PamtraFortranWrapper

missingNumber = -9999.
missingIntNumber = int(missingNumber)


class pyPamtra(object):
    """Settings, profile and results of a forward-operator run
    (python/pyPamtra/core.py:74-274)."""

    def __init__(self):
        self.default_p_vars = [
            "timestamp", "lat", "lon", "wind10u", "wind10v", "hgt", "press", "temp", "relhum",
            "hgt_lev", "press_lev", "temp_lev", "relhum_lev", "q", "hydro_q", "hydro_n",
            "hydro_reff", "obs_height", "ngridy", "ngridx", "max_nlyrs", "nlyrs", "model_i",
            "model_j", "unixtime", "airturb", "radar_prop", "groundtemp", "wind_w", "wind_uv",
            "turb_edr", "sfc_type", "sfc_model", "sfc_refl", "sfc_salinity", "sfc_slf",
            "sfc_sif", "sfc_emissivity"]
        n = dict()
        n["hydro_threshold"] = 1.e-10
        n["write_nc"] = True
        n["data_path"] = '$PAMTRA_DATADIR'
        n["save_psd"] = False
        n["save_ssp"] = False
        n["add_obs_height_to_layer"] = False
        n["conserve_mass_rescale_dsd"] = True
        n["outpol"] = 'VH'
        n["file_desc"] = ''
        n["creator"] = 'Pamtrauser'
        n["active"] = True
        n["passive"] = True
        n["radar_mode"] = "simple"
        n["randomseed"] = 0
        n["emissivity"] = 0.6
        n["lgas_extinction"] = True
        n["gas_mod"] = 'R98'
        n["lhyd_absorption"] = True
        n["lhyd_scattering"] = True
        n["lhyd_emission"] = True
        n["liq_mod"] = "TKC"
        n["hydro_includehydroinrhoair"] = True
        n["hydro_fullspec"] = False
        n["hydro_limit_density_area"] = True
        n["hydro_softsphere_min_density"] = 10.
        n["hydro_adaptive_grid"] = True
        n["tmatrix_db"] = "none"
        n["tmatrix_db_path"] = "database/"
        n["radar_nfft"] = 256
        n["radar_no_ave"] = 150
        n["radar_max_v"] = 7.885
        n["radar_min_v"] = -7.885
        n["radar_pnoise0"] = -32.23
        n["radar_allow_negative_dD_dU"] = False
        n["radar_airmotion"] = False
        n["radar_airmotion_model"] = "constant"
        n["radar_airmotion_vmin"] = -4.e0
        n["radar_airmotion_vmax"] = +4.e0
        n["radar_airmotion_linear_steps"] = 30
        n["radar_airmotion_step_vmin"] = 0.5e0
        n["radar_peak_min_bins"] = 2
        n["radar_aliasing_nyquist_interv"] = 1
        n["radar_save_noise_corrected_spectra"] = False
        n["radar_use_hildebrand"] = False
        n["radar_peak_snr_definition"] = 'log'
        n["radar_peak_min_snr"] = -10
        n["radar_convolution_fft"] = True
        n["radar_smooth_spectrum"] = True
        n["radar_k2"] = 0.93
        n["radar_npeaks"] = 1
        n["radar_noise_distance_factor"] = 2.0
        n["radar_receiver_uncertainty_std"] = 0.e0
        n["radar_receiver_miscalibration"] = 0.e0
        n["radar_attenuation"] = "disabled"
        n["radar_polarisation"] = "NN"
        n["radar_use_wider_peak"] = False
        n["radar_integration_time"] = 1.4
        n["radar_fwhr_beamwidth_deg"] = 0.31
        n["liblapack"] = True
        self.nmlSet = n

        self.set = dict()
        self.set["pyVerbose"] = 0
        self.set["verbose"] = 0
        self.set["freqs"] = []
        self.set["namelist_file"] = "TMPFILE"
        # extension: arithmetic of the device path, "fp64" or "mixed"
        self.set["precision"] = "fp64"
        self._nmlDefaultSet = deepcopy(self.nmlSet)

        self.dimensions = {
            "tb": ["gridx", "gridy", "outlevels", "angles", "frequency", "passive_npol"],
            "hgt_lev": ["ngridx", "ngridy", "max_nlyrs+1"],
            "hydro_q": ["ngridx", "ngridy", "max_nlyrs", "nhydro"],
        }
        self.units = {"tb": "K", "hgt_lev": "m", "temp_lev": "K", "p_lev": "Pa",
                      "relhum_lev": "%", "hydro_q": "kg/kg"}
        self._nstokes = 2
        self._nangles = 16
        self.p = dict()
        self.r = dict()
        self.p["ngridx"] = 0
        self.p["ngridy"] = 0
        self.p["max_nlyrs"] = 0
        self._shape2D = (0, 0)
        self._shape3D = (0, 0, 0)
        self._shape4D = (0, 0, 0, 0)
        self.df = pamDescriptorFile(self)

    # ------------------------------------------------------------------
    def createProfile(self, **kwargs):
        """Builds ``self.p`` from level/layer fields (core.py:756-976).
        hgt(_lev), temp(_lev), press(_lev) and relhum(_lev) are mandatory;
        everything else gets the reference's defaults."""
        for key in list(kwargs.keys()):
            v = kwargs[key]
            if isinstance(v, np.ma.MaskedArray):
                v = v.filled(np.nan)
            if isinstance(v, np.char.chararray):
                continue
            kwargs[key] = np.array(v)

        if not self.df.data.shape[0] > 0:
            warnings.warn("No descriptor file defined. Assuming dry profile without hydrometeors")
        for qv in ("hydro_q", "hydro_n", "hydro_reff"):
            if qv in kwargs:
                assert self.df.nhydro == kwargs[qv].shape[-1]

        for key in kwargs:
            if (key not in self.default_p_vars and key.split("+")[0] not in self.default_p_vars) \
                    or key in ("model_i", "model_j"):
                raise TypeError("Could not parse " + key)
        if not ((("hgt_lev" in kwargs) or ("hgt" in kwargs))
                and (("temp_lev" in kwargs) or ("temp" in kwargs))
                and (("press_lev" in kwargs) or ("press" in kwargs))
                and (("relhum_lev" in kwargs) or ("relhum" in kwargs))):
            raise TypeError("I need hgt(_lev) and temp(_lev) and press(_lev) and relhum(_lev)!")

        if "hgt" not in kwargs:
            kwargs["hgt"] = (kwargs["hgt_lev"][..., 1:] + kwargs["hgt_lev"][..., :-1]) / 2.

        nd = kwargs["hgt"].ndim
        if nd == 1:
            self.p["ngridx"], self.p["ngridy"] = 1, 1
        elif nd == 2:
            self.p["ngridx"], self.p["ngridy"] = kwargs["hgt"].shape[0], 1
        elif nd == 3:
            self.p["ngridx"], self.p["ngridy"] = kwargs["hgt"].shape[0], kwargs["hgt"].shape[1]
        else:
            raise IOError("Too many dimensions!")
        self.p["max_nlyrs"] = kwargs["hgt"].shape[-1]
        self.p["nlyrs"] = np.array(np.sum(kwargs["hgt"] != missingNumber, axis=-1))
        try:
            self.p["noutlevels"] = np.shape(kwargs["obs_height"])[-1]
        except (KeyError, IndexError):
            self.p["noutlevels"] = 2

        nx, ny, L = self.p["ngridx"], self.p["ngridy"], self.p["max_nlyrs"]
        nh = self.df.nhydro
        self._shape2D = (nx, ny)
        self._shape3D = (nx, ny, L)
        self._shape3Dout = (nx, ny, self.p["noutlevels"])
        self._shape3Dplus = (nx, ny, L + 1)
        self._shape4D = (nx, ny, L, nh)
        self.p["nlyrs"] = self.p["nlyrs"].reshape(self._shape2D)

        for key in ("hgt_lev", "temp_lev", "press_lev", "relhum_lev"):
            if key in kwargs:
                self.p[key] = kwargs[key].reshape(self._shape3Dplus)
        for key in ("hgt", "temp", "press", "relhum"):
            if key in kwargs:
                self.p[key] = kwargs[key].reshape(self._shape3D)
        if "hgt_lev" not in kwargs:
            h = self.p["hgt"]
            hl = np.empty(self._shape3Dplus)
            hl[..., 0] = h[..., 0] - (h[..., 1] - h[..., 0]) * 0.5
            hl[..., 1:-1] = 0.5 * (h[..., :-1] + h[..., 1:])
            hl[..., -1] = h[..., -1] + (h[..., -1] - h[..., -2]) * 0.5
            self.p["hgt_lev"] = hl

        # 1-based grid indices of the columns (core.py:885-886)
        gi, gj = np.indices(self._shape2D)
        self.p["model_i"], self.p["model_j"] = gi + 1, gj + 1

        if "timestamp" not in kwargs:
            self.p["unixtime"] = np.ones(self._shape2D) * int(time.time())
        else:
            self.p["unixtime"] = np.ones(self._shape2D) * np.asarray(kwargs["timestamp"])

        for env, preset in (("lat", 50.938056), ("lon", 6.956944), ("wind10u", 0),
                            ("wind10v", 0), ("groundtemp", np.nan), ("sfc_salinity", 33.),
                            ("sfc_slf", 1.), ("sfc_sif", 0.)):
            if env not in kwargs:
                self.p[env] = np.ones(self._shape2D) * preset
            else:
                self.p[env] = (np.ones(self._shape2D) * kwargs[env]) if kwargs[env].ndim == 0 \
                    else kwargs[env].reshape(self._shape2D)
        for env, preset in (("sfc_type", -9999), ("sfc_model", -9999)):
            if env not in kwargs:
                self.p[env] = np.ones(self._shape2D, dtype=int) * preset
            else:
                self.p[env] = (np.ones(self._shape2D) * kwargs[env]) if kwargs[env].ndim == 0 \
                    else kwargs[env].reshape(self._shape2D)
        if "sfc_refl" not in kwargs:
            self.p["sfc_refl"] = np.full(self._shape2D, "S", dtype="U1")
        else:
            r = np.asarray(kwargs["sfc_refl"]).astype("U1")
            self.p["sfc_refl"] = np.full(self._shape2D, str(r), dtype="U1") if r.ndim == 0 \
                else r.reshape(self._shape2D)
        if "sfc_emissivity" in kwargs:
            self.p["sfc_emissivity"] = kwargs["sfc_emissivity"].reshape(self._shape2D)
        if "obs_height" not in kwargs:
            self.p["obs_height"] = np.ones(self._shape3Dout) * np.array([833000., 0.])
        else:
            oh = kwargs["obs_height"]
            self.p["obs_height"] = (np.ones(self._shape3Dout) * oh) if oh.ndim <= 1 \
                else oh.reshape(self._shape3Dout)
        for qv in ("hydro_q", "hydro_reff", "hydro_n"):
            if qv not in kwargs and qv + "+no000" not in kwargs:
                self.p[qv] = np.zeros(self._shape4D)
            elif qv + "+no000" in kwargs:
                self.p[qv] = np.zeros(self._shape4D)
                for hh in range(nh):
                    self.p[qv][..., hh] = kwargs[qv + "+no" + str(hh).zfill(3)].reshape(self._shape3D)
            else:
                self.p[qv] = kwargs[qv].reshape(self._shape4D)
        for qv in ("airturb", "wind_w", "wind_uv", "turb_edr"):
            self.p[qv] = kwargs[qv].reshape(self._shape3D) if qv in kwargs \
                else np.zeros(self._shape3D) + np.nan

    # ------------------------------------------------------------------
    # ASCII profile files (core.py:339-504, 651-749).  File layout, per
    # grid point: "year month day hhmm nheights i j", the observation
    # heights, "lat lon lfrac wind10u wind10v groundtemp hgt_lev0", the
    # integrated values, then one line per layer (.lay) or level (.lev, with
    # the ground level first): height [m], pressure [Pa], temperature [K],
    # relative humidity [%] and the hydrometeor moments named by each
    # category's moment_in.
    _MOMENT_KEYS = {0: (), 1: ("hydro_n",), 2: ("hydro_reff",), 3: ("hydro_q",),
                    12: ("hydro_n", "hydro_reff"), 13: ("hydro_n", "hydro_q"),
                    23: ("hydro_reff", "hydro_q")}

    def _moment_keys(self, hh):
        mi = int(self.df.data["moment_in"][hh])
        if mi not in self._MOMENT_KEYS:
            raise IOError('Did not understand df.data["moment_in"]')
        return self._MOMENT_KEYS[mi]

    def readPamtraProfile(self, inputFile):
        """Reads a .lay or .lev profile (core.py:339-504); the descriptor
        file must be defined first.  Any other extension is read as the
        classic format (readClassicPamtraProfile)."""
        import calendar
        import datetime
        assert self.df.nhydro > 0
        kind = inputFile.split(".")[-1]
        if kind not in ("lay", "lev"):
            self.readClassicPamtraProfile(inputFile)
            return
        with open(inputFile) as fh:
            rows = (line.split() for line in fh if line.strip())
            nx, ny, L, nout = (int(v) for v in next(rows)[:4])
            self.p["ngridx"], self.p["ngridy"], self.p["max_nlyrs"], self.p["noutlevels"] = \
                nx, ny, L, nout
            nh = self.df.nhydro
            self._shape2D = (nx, ny)
            self._shape3D = (nx, ny, L)
            self._shape3Dplus = (nx, ny, L + 1)
            self._shape3Dout = (nx, ny, nout)
            self._shape4D = (nx, ny, L, nh)
            p = self.p
            for k in ("lat", "lon", "wind10u", "wind10v", "groundtemp", "sfc_salinity", "iwv",
                      "sfc_slf", "sfc_sif"):
                p[k] = np.full(self._shape2D, np.nan)
            for k in ("model_i", "model_j", "unixtime", "nlyrs", "sfc_type", "sfc_model"):
                p[k] = np.full(self._shape2D, missingIntNumber, dtype=int)
            p["sfc_refl"] = np.full(self._shape2D, "S", dtype="U1")
            p["obs_height"] = np.full(self._shape3Dout, np.nan)
            for k in ("hydro_q", "hydro_n", "hydro_reff"):
                p[k] = np.full(self._shape4D, np.nan)
            for k in ("hgt", "temp", "press", "relhum", "airturb", "wind_w", "wind_uv",
                      "turb_edr"):
                p[k] = np.full(self._shape3D, np.nan)
            for k in ("hgt_lev", "temp_lev", "press_lev", "relhum_lev"):
                p[k] = np.full(self._shape3Dplus, np.nan)
            for x in range(nx):
                for y in range(ny):
                    year, mon, day, hhmm, nhgt, mi, mj = next(rows)[:7]
                    p["unixtime"][x, y] = calendar.timegm(datetime.datetime(
                        int(year), int(mon), int(day), int(hhmm[0:2]), int(hhmm[2:4])).timetuple())
                    p["model_i"][x, y], p["model_j"][x, y] = int(mi), int(mj)
                    p["obs_height"][x, y, :] = [float(v) for v in next(rows)[:nout]]
                    lat, lon, lfrac, u, v, tg, h0 = (float(t) for t in next(rows)[:7])
                    p["lat"][x, y], p["lon"][x, y] = lat, lon
                    p["wind10u"][x, y], p["wind10v"][x, y], p["groundtemp"][x, y] = u, v, tg
                    p["hgt_lev"][x, y, 0] = h0
                    # lfrac (deprecated) -> surface type: 0 ocean (Fresnel), else land
                    p["sfc_type"][x, y] = int(np.around(lfrac))
                    if p["sfc_type"][x, y] == 0:
                        p["sfc_refl"][x, y] = "F"
                        p["sfc_salinity"][x, y] = 33.0
                    p["iwv"][x, y] = float(next(rows)[0])
                    n = int(nhgt)
                    if kind == "lev":
                        vals = [float(t) for t in next(rows)]
                        (p["hgt_lev"][x, y, 0], p["press_lev"][x, y, 0], p["temp_lev"][x, y, 0],
                         p["relhum_lev"][x, y, 0]) = vals[:4]
                        n -= 1  # the file counts levels
                    p["nlyrs"][x, y] = n
                    for z in range(n):
                        vals = [float(t) for t in next(rows)]
                        if kind == "lay":
                            (p["hgt"][x, y, z], p["press"][x, y, z], p["temp"][x, y, z],
                             p["relhum"][x, y, z]) = vals[:4]
                        else:
                            (p["hgt_lev"][x, y, z + 1], p["press_lev"][x, y, z + 1],
                             p["temp_lev"][x, y, z + 1], p["relhum_lev"][x, y, z + 1]) = vals[:4]
                        k = 4
                        for hh in range(nh):
                            for key in self._moment_keys(hh):
                                p[key][x, y, z, hh] = vals[k]
                                k += 1
                        if k != len(vals):
                            raise IOError(f"{inputFile}: {len(vals)} values on a line, {k} expected")
        if kind == "lay":
            p["hgt_lev"][..., 1:-1] = 0.5 * (p["hgt"][..., :-1] + p["hgt"][..., 1:])
            p["hgt_lev"][..., -1] = p["hgt"][..., -1] + (p["hgt"][..., -1] - p["hgt"][..., -2]) * 0.5

    def readClassicPamtraProfile(self, inputFile, n_moments=1):
        """Reads the classic profile format (core.py:505-649): a header
        "year month day hhmm ngridx ngridy nlyrs deltax deltay", then per
        column the model indices, "lat lon lfrac wind10u wind10v", the paths
        "iwv cwp iwp rwp swp gwp [hwp]", the ground level "hgt press temp
        relhum[%]" and one line per layer with the upper level and the
        contents of cloud, ice, rain, snow, graupel (and, with n_moments = 2,
        hail and the six number concentrations).  The categories become
        hydro_q / hydro_n in that order (4 + n_moments of them)."""
        import calendar
        import datetime
        if n_moments not in (1, 2):
            raise IOError("n_moments mus be 1 or 2")
        with open(inputFile) as fh:
            rows = (line.split() for line in fh if line.strip())
            year, mon, day, hhmm, nx, ny, L = next(rows)[:7]
            nx, ny, L = int(nx), int(ny), int(L)
            p = self.p
            p["ngridx"], p["ngridy"], p["max_nlyrs"], p["noutlevels"] = nx, ny, L, 2
            nh = 4 + n_moments
            self._shape2D = (nx, ny)
            self._shape3D = (nx, ny, L)
            self._shape3Dplus = (nx, ny, L + 1)
            self._shape3Dout = (nx, ny, 2)
            self._shape4D = (nx, ny, L, nh)
            for k in ("model_i", "model_j", "lat", "lon", "wind10u", "wind10v", "iwv", "cwp",
                      "iwp", "rwp", "swp", "gwp", "hwp"):
                p[k] = np.zeros(self._shape2D)
            p["sfc_type"] = np.full(self._shape2D, missingIntNumber, dtype=int)
            p["radar_prop"] = np.full(self._shape2D + (2,), np.nan)
            for k in ("hgt_lev", "temp_lev", "press_lev", "relhum_lev"):
                p[k] = np.zeros(self._shape3Dplus)
            q = np.zeros(self._shape4D)
            n = np.zeros(self._shape4D)
            paths = ("iwv", "cwp", "iwp", "rwp", "swp", "gwp", "hwp")[:6 + n_moments - 1]
            for x in range(nx):
                for y in range(ny):
                    p["model_i"][x, y], p["model_j"][x, y] = (int(v) for v in next(rows))
                    lat, lon, lfrac, u, v = (float(t) for t in next(rows))
                    p["lat"][x, y], p["lon"][x, y] = lat, lon
                    p["wind10u"][x, y], p["wind10v"][x, y] = u, v
                    for key, val in zip(paths, next(rows), strict=True):
                        p[key][x, y] = float(val)
                    (p["hgt_lev"][x, y, 0], p["press_lev"][x, y, 0], p["temp_lev"][x, y, 0],
                     p["relhum_lev"][x, y, 0]) = (float(t) for t in next(rows))
                    p["sfc_type"][x, y] = np.around(lfrac)  # lfrac is deprecated
                    for z in range(L):
                        vals = [float(t) for t in next(rows)]
                        if len(vals) != 4 + nh + (nh if n_moments == 2 else 0):
                            raise IOError(f"{inputFile}: {len(vals)} values on a layer line")
                        (p["hgt_lev"][x, y, z + 1], p["press_lev"][x, y, z + 1],
                         p["temp_lev"][x, y, z + 1], p["relhum_lev"][x, y, z + 1]) = vals[:4]
                        q[x, y, z, :] = vals[4:4 + nh] if n_moments == 2 else vals[4:]
                        if n_moments == 2:
                            n[x, y, z, :] = vals[4 + nh:]
        p["relhum_lev"] = p["relhum_lev"] / 100.
        p["nlyrs"] = np.full(self._shape2D, L, dtype=int)
        p["unixtime"] = np.full(self._shape2D, calendar.timegm(datetime.datetime(
            int(year), int(mon), int(day), int(hhmm[0:2]), int(hhmm[2:4])).timetuple()), dtype=int)
        p["hydro_q"], p["hydro_n"], p["hydro_reff"] = q, n, np.zeros(self._shape4D)
        for k in ("airturb", "wind_w", "wind_uv", "turb_edr"):
            p[k] = np.full(self._shape3D, np.nan)

    def writePamtraProfile(self, profileFile):
        """Writes the profile as .lay or .lev, by the file extension
        (core.py:651-749), with the reference's number formats."""
        import datetime
        kind = profileFile.split(".")[-1]
        if kind not in ("lay", "lev"):
            raise IOError("Did not understand lay/lev: " + kind)
        p = self.p
        nx, ny, L = self._shape3D
        nout = self._shape3Dout[2]
        t0 = datetime.datetime.fromtimestamp(int(np.ravel(p["unixtime"])[0]), datetime.timezone.utc)
        stamp = f"{t0.year} {t0.month:02d} {t0.day:02d} {t0:%H%M}"
        iwv = p.get("iwv", np.full(self._shape2D, missingNumber))
        out = [f"{nx} {ny} {L} {nout}"]
        for x in range(nx):
            for y in range(ny):
                out.append(f"{stamp} {L + 1 if kind == 'lev' else L} {x + 1} {y + 1}")
                out.append(" ".join("%9e" % h for h in p["obs_height"][x, y, :]))
                out.append("%3.2f %3.2f %s %s %s %s %s" % (
                    p["lat"][x, y], p["lon"][x, y], p["sfc_type"][x, y], p["wind10u"][x, y],
                    p["wind10v"][x, y], p["groundtemp"][x, y], p["hgt_lev"][x, y, 0]))
                line = str(iwv[x, y] if np.isfinite(iwv[x, y]) else missingNumber)
                for hh in range(self.df.nhydro):
                    line += "".join(" %9e" % missingNumber for _ in self._moment_keys(hh))
                out.append(line)
                if kind == "lev":
                    out.append("%6.1f %6.1f %3.2f %1.4f" % (
                        p["hgt_lev"][x, y, 0], p["press_lev"][x, y, 0], p["temp_lev"][x, y, 0],
                        p["relhum_lev"][x, y, 0]))
                for z in range(L):
                    if kind == "lev":
                        f = (p["hgt_lev"][x, y, z + 1], p["press_lev"][x, y, z + 1],
                             p["temp_lev"][x, y, z + 1], p["relhum_lev"][x, y, z + 1])
                    else:
                        f = (p["hgt"][x, y, z], p["press"][x, y, z], p["temp"][x, y, z],
                             p["relhum"][x, y, z])
                    line = "%6.1f %6.1f %3.2f %1.4f " % f
                    for hh in range(self.df.nhydro):
                        line += "".join("%9e " % p[key][x, y, z, hh] for key in self._moment_keys(hh))
                    out.append(line)
        with open(profileFile, "w") as fh:
            fh.write("\n".join(out) + "\n")

    # ------------------------------------------------------------------
    def rescaleHeights(self, new_hgt_lev, new_hgt=[]):
        """Moves the profile to a new height grid (core.py:1161-1288): every
        level and layer field is interpolated linearly in height, pressure
        in its logarithm, and continued linearly beyond the old grid;
        negative hydrometeor moments become 0, other negative values
        missing.  wind_w, which the reference leaves out, is reset to
        missing on the new grid."""
        assert len(self.df.dataFullSpec) == 0, \
            "rescaleHeights does not regrid the full-spectrum input (hydro_fullspec)"

        def extrap(x, xp, yp):
            y = np.interp(x, xp, yp)
            lo, hi = x < xp[0], x > xp[-1]
            y[lo] = yp[0] + (x[lo] - xp[0]) * (yp[0] - yp[1]) / (xp[0] - xp[1])
            y[hi] = yp[-1] + (x[hi] - xp[-1]) * (yp[-1] - yp[-2]) / (xp[-1] - xp[-2])
            return y

        new_hgt_lev = np.asarray(new_hgt_lev, dtype=float)
        old_hgt_lev, old_hgt = self.p["hgt_lev"], self.p["hgt"]
        if len(new_hgt) == 0:
            new_hgt = 0.5 * (new_hgt_lev[..., 1:] + new_hgt_lev[..., :-1])
        nx, ny = self._shape2D
        L = new_hgt_lev.shape[2] - 1
        nh = self.df.nhydro
        self.p["max_nlyrs"] = L
        self._shape3D = (nx, ny, L)
        self._shape3Dplus = (nx, ny, L + 1)
        self._shape4D = (nx, ny, L, nh)

        def regrid(field, new_z, old_z, shape, log=False):
            out = np.full(shape, missingNumber)
            for x in range(nx):
                for y in range(ny):
                    if len(shape) == 4:
                        for h in range(nh):
                            out[x, y, :, h] = extrap(new_z[x, y], old_z[x, y], field[x, y, :, h])
                    elif log:
                        out[x, y] = np.exp(extrap(new_z[x, y], old_z[x, y], np.log(field[x, y])))
                    else:
                        out[x, y] = extrap(new_z[x, y], old_z[x, y], field[x, y])
            return out

        for key in ("hydro_q", "hydro_n", "hydro_reff"):
            if key in self.p:
                self.p[key] = regrid(self.p[key], new_hgt, old_hgt, self._shape4D)
                self.p[key][self.p[key] < 0.0] = 0.0
        for key in list(self.df.data4D):
            self.df.data4D[key] = regrid(self.df.data4D[key], new_hgt, old_hgt, self._shape4D)
        for key in ("hgt_lev", "temp_lev", "relhum_lev"):
            if key in self.p:
                self.p[key] = regrid(self.p[key], new_hgt_lev, old_hgt_lev, self._shape3Dplus)
                if key != "hgt_lev":
                    self.p[key][self.p[key] < -1.e-6] = missingNumber
        for key in ("airturb", "hgt", "temp", "relhum", "wind_uv", "turb_edr"):
            if key in self.p:
                self.p[key] = regrid(self.p[key], new_hgt, old_hgt, self._shape3D)
                if key != "hgt":
                    self.p[key][self.p[key] < -1.e-6] = missingNumber
        if "press_lev" in self.p:
            self.p["press_lev"] = regrid(self.p["press_lev"], new_hgt_lev, old_hgt_lev,
                                         self._shape3Dplus, log=True)
            self.p["press_lev"][self.p["press_lev"] < -1.e-6] = missingNumber
        if "press" in self.p:
            self.p["press"] = regrid(self.p["press"], new_hgt, old_hgt, self._shape3D, log=True)
            self.p["press"][self.p["press"] < -1.e-6] = missingNumber
        if "wind_w" in self.p:
            self.p["wind_w"] = np.full(self._shape3D, np.nan)
        self.p["nlyrs"] = np.sum(self.p["hgt_lev"] != missingNumber, axis=-1) - 1

    def _addObservationLevels(self):
        """Adds the observation heights to the level grid of every column
        (core.py:1908-1951), then rescales the profile.  As in the
        reference, the first observation level is taken as the top and the
        last as the ground when all columns have them there, and a height
        on an existing level, at or above the top, or at or below the ground
        is moved 1 m into the column."""
        oh = self.p["obs_height"]
        top, ground = self.p["hgt_lev"][:, :, -1], self.p["hgt_lev"][:, :, 0]
        nobs = oh.shape[2]
        skip_top = 1 if any(np.all(oh[:, :, i] >= top) for i in range(nobs)) else 0
        skip_bot = 1 if any(np.all(oh[:, :, i] <= ground) for i in range(nobs)) else 0
        nx, ny, nl = self._shape3Dplus
        new = np.full((nx, ny, nl + nobs - skip_top - skip_bot), np.nan)
        for x in range(nx):
            for y in range(ny):
                lev = self.p["hgt_lev"][x, y, :]
                for k in range(skip_top, nobs - skip_bot):
                    h = oh[x, y, k]
                    if np.any(h == lev[1:-1]):
                        h = h + 1.0
                    elif h >= lev[-1]:
                        h = lev[-1] - 1.0
                    elif h <= lev[0]:
                        h = lev[0] + 1.0
                    lev = np.sort(np.append(lev, h))
                new[x, y, :] = lev
        self.rescaleHeights(new)

    # ------------------------------------------------------------------
    def _checkData(self):
        """Range checks of core.py:979-999."""
        maxLimits = {"relhum_lev": 200, "hydro_q": 0.05, "temp_lev": 320, "press_lev": 110000,
                     "relhum": 200, "temp": 320, "press": 110000}
        minLimits = {"relhum_lev": 0, "hydro_q": 0, "temp_lev": 160, "press_lev": 1,
                     "relhum": 0, "temp": 160, "press": 1}
        for key, v in self.p.items():
            if not isinstance(v, np.ndarray) or key == "sfc_refl" or v.dtype.kind not in "fiu":
                continue
            d = v[~np.isnan(v)]
            if key in maxLimits and len(d) > 0 and np.max(d) > maxLimits[key]:
                raise ValueError("unrealistic value for " + key + ": " + str(np.max(d))
                                 + ", maximum is " + str(maxLimits[key]))
            if key in minLimits and len(d) > 0:
                dd = d[d != missingNumber]
                if len(dd) > 0 and np.min(dd) < minLimits[key]:
                    raise ValueError("unrealistic value for " + key + ": " + str(np.min(d))
                                     + ", minimum is " + str(minLimits[key]))
        if np.any(self.p["sfc_refl"][self.p["sfc_type"] == 0] == "L"):
            raise ValueError("It does not make sense to treat Ocean as Lambertian reflector "
                             "(incompatible sfc_refl and sfc_type).")
        if np.any(self.p["sfc_refl"][self.p["sfc_type"] == 1] == "F"):
            raise ValueError("It does not make sense to simulate land with Fresnel reflection "
                             "(incompatible sfc_refl and sfc_type).")

    # ------------------------------------------------------------------
    def runPamtra(self, freqs, checkData=True, failOnError=True):
        """Runs the forward operator and fills ``self.r`` (core.py:1518-1568).
        freqs [GHz] must be sorted ascending."""
        tttt = time.time()
        if isinstance(freqs, (int, float, np.integer, np.floating)):
            freqs = [freqs]
        assert np.array_equal(np.sort(freqs), freqs)
        self.set["freqs"] = freqs
        self.set["nfreqs"] = len(freqs)
        self.set["radar_pol"] = self.nmlSet["radar_polarisation"].split(",")
        self.set["radar_npol"] = len(self.set["radar_pol"])
        self.set["att_pol"] = ["N"]
        self.set["att_npol"] = 1
        assert self.set["nfreqs"] > 0
        assert self.p["noutlevels"] > 0
        assert np.prod(self._shape2D) > 0
        if checkData:
            self._checkData()
        if self.nmlSet["add_obs_height_to_layer"]:
            self._addObservationLevels()
        # the angle count is the reference's compile-time nummu (settings.f90:16),
        # mirrored as _nangles: 16, or 32 for the 32-angle library
        sets = dict(self.set, _nangles=self._nangles)
        fortResults, self.fortError, fortObject = PamtraFortranWrapper(
            sets, self.nmlSet, self.df.data, self.df.data4D, self.df.dataFullSpec, self.p)
        if failOnError and self.fortError != 0:
            raise RuntimeError("Fortran code failed")
        self.r = fortResults
        self.fortObject = fortObject
        self.r["nmlSettings"] = self.nmlSet
        if self.set["pyVerbose"] > 0:
            print("pyPamtra runtime:", time.time() - tttt)

    # ------------------------------------------------------------------
    def writeResultsToNetCDF(self, fname, profileVars="all", wpNames=[], ncForm="NETCDF3_CLASSIC",
                             xarrayCompatibleOutput=False, ncCompression=False,
                             lfracCompatibility=False):
        """Writes the results to a netCDF file with the reference's dimensions,
        variables, types, units, fill values and global attributes
        (core.py:2163-2556).

        The netCDF-3 forms are written by scipy's netCDF-3 writer:
        NETCDF3_CLASSIC and NETCDF3 as the classic format, NETCDF3_64BIT with
        64-bit offsets.  The HDF5-based NETCDF4 and NETCDF4_CLASSIC forms need
        the netCDF4 package.  The single-character passive_polarisation
        variable carries no _FillValue.  Radar spectra and moments are not
        produced by this library, so their variables are never written."""
        import datetime

        from scipy.io import netcdf_file

        forms = {"NETCDF3_CLASSIC": 1, "NETCDF3": 1, "NETCDF3_64BIT": 2}
        if ncForm in ("NETCDF4_CLASSIC", "NETCDF4"):
            raise NotImplementedError(f"ncForm {ncForm!r} needs the netCDF4 package; "
                                      "use NETCDF3_CLASSIC or NETCDF3_64BIT")
        if ncForm not in forms:
            raise ValueError("Unknown nc form " + ncForm)
        if not hasattr(self, "r") or "pamtraVersion" not in self.r:
            raise IOError("run runPamtra first!")
        nml = self.r["nmlSettings"]
        f = netcdf_file(fname, "w", version=forms[ncForm])
        try:
            f.history = ("Created with pyPamtra (Version: " + self.r["pamtraVersion"] +
                         ", Git Hash: " + self.r["pamtraHash"] + ") by " + self.nmlSet["creator"] +
                         " (University of Cologne, IGMK) at " +
                         datetime.datetime.now().strftime("%Y/%m/%d %H:%M:%S"))
            f.properties = str(nml)
            f.createDimension("grid_x", int(self.p["ngridx"]))
            f.createDimension("grid_y", int(self.p["ngridy"]))
            f.createDimension("frequency", int(self.set["nfreqs"]))
            outdim = "outlevel" if xarrayCompatibleOutput else "outlevels"
            if nml["passive"]:
                f.createDimension("angles", len(self.r["angles_deg"]))
                f.createDimension(outdim, int(self.p["noutlevels"]))
                f.createDimension("passive_polarisation", int(self._nstokes))
            if nml["active"]:
                f.createDimension("heightbins", int(self.p["max_nlyrs"]))
                f.createDimension("radar_polarisation", int(self.set["radar_npol"]))
                f.createDimension("attenuation_polarisation", int(self.set["att_npol"]))
                f.createDimension("radar_peak_number", int(self.nmlSet["radar_npeaks"]))
            dim2d = ("grid_x", "grid_y")

            def var(name, typ, dims, units, data):
                v = f.createVariable(name, typ, dims)
                v.units = units
                dt = np.float32 if typ == "f" else np.int32
                v._FillValue = dt(missingNumber)
                v[:] = np.asarray(data, dtype=dt).reshape(v.shape)
                return v

            var("frequency", "f", ("frequency",), "GHz", self.set["freqs"])
            var("grid_x", "f", ("grid_x",), "-", np.arange(self.p["ngridx"]))
            var("grid_y", "f", ("grid_y",), "-", np.arange(self.p["ngridy"]))
            if nml["active"]:
                var("heightbins", "i", ("heightbins",), "-", np.arange(self.p["max_nlyrs"]))
                var("height", "f", dim2d + ("heightbins",), "m", self.r["radar_hgt"])
            if nml["passive"]:
                var("angles", "f", ("angles",), "deg", self.r["angles_deg"])
                v = f.createVariable("passive_polarisation", "c", ("passive_polarisation",))
                v.units = "-"
                v[:] = np.array(["V", "H"][:self._nstokes], dtype="S1")
                if xarrayCompatibleOutput:
                    var("outlevel", "f", (outdim,), "-", np.arange(self.p["noutlevels"]))
                var("outlevels", "f", dim2d + (outdim,), "m over sea level",
                    self.p["obs_height"])
            var("model_i", "i", dim2d, "-", self.p["model_i"])
            var("model_j", "i", dim2d, "-", self.p["model_j"])
            var("nlyr", "i", dim2d, "-", self.p["nlyrs"])
            var("datatime", "i", dim2d, "seconds since 1970-01-01 00:00:00", self.p["unixtime"])
            var("longitude", "f", dim2d, "deg.dec", self.p["lon"])
            var("latitude", "f", dim2d, "deg.dec", self.p["lat"])
            if lfracCompatibility:
                var("lfrac", "f", dim2d, "-", self.p["sfc_type"])
            var("sfc_type", "i", dim2d, "0: water, 1: land", self.p["sfc_type"])
            if nml["active"]:
                d3 = dim2d + ("heightbins",)
                var("Ze", "f", d3 + ("frequency", "radar_polarisation", "radar_peak_number"),
                    "dBz", self.r["Ze"])
                var("Attenuation_Hydrometeors", "f", d3 + ("frequency", "attenuation_polarisation"),
                    "dB", self.r["Att_hydro"])
                var("Attenuation_Atmosphere", "f", d3 + ("frequency",), "dB", self.r["Att_atmo"])
                if "PIA" in self.r:
                    var("Path_Integrated_Attenuation", "f", d3 + ("frequency",), "dB",
                        self.r["PIA"])
            if nml["passive"]:
                var("tb", "f", dim2d + (outdim, "angles", "frequency", "passive_polarisation"), "K",
                    self.r["tb"])
            if (profileVars == "all" or "iwv" in profileVars) and "iwv" in self.p:
                var("iwv", "f", dim2d, "kg/m^2", self.p["iwv"])
            if (profileVars == "all" or "hydro_wp" in profileVars) and "hydro_wp" in self.p:
                for i, wp in enumerate(wpNames):
                    var(wp, "f", dim2d, "kg/m^2", self.p["hydro_wp"][..., i])
            for key in ("cloudBase", "cloudTop"):
                if (profileVars == "all" or key in profileVars) and key in self.p:
                    var(key, "f", dim2d, "m", self.p[key])
        finally:
            f.close()
        if self.set["pyVerbose"] > 0:
            print(fname, "written")

    # ------------------------------------------------------------------
    # profile handling and session files around the run (core.py:277-338,
    # 1001-1160, 1337-1358, 1570-1675, 2072-2162, 2555-2594)

    def _set_shapes(self):
        p, nh = self.p, self.df.nhydro
        nx, ny, L = p["ngridx"], p["ngridy"], p["max_nlyrs"]
        self.df.fs_nbin = self.df.dataFullSpec["d_ds"].shape[-1] if "d_ds" in self.df.dataFullSpec \
            else 0
        self._shape2D = (nx, ny)
        self._shape3D = (nx, ny, L)
        self._shape3Dplus = (nx, ny, L + 1)
        self._shape3Dout = (nx, ny, p["noutlevels"])
        self._shape3Dhyd = (nx, ny, nh)
        self._shape4D = (nx, ny, L, nh)
        self._shape5Dplus = (nx, ny, L, nh, self.df.fs_nbin + 1)
        self._shape5D = (nx, ny, L, nh, self.df.fs_nbin)

    def _reshape_all(self, take):
        """Applies take(array, trailing_ndim) to every per-column array of
        the profile, the 4D descriptor and the full spectrum."""
        p = self.p
        for k in list(p):
            v = p[k]
            if isinstance(v, np.ndarray) and v.ndim >= 2:
                p[k] = take(v, v.ndim - 2)
        for k in list(self.df.data4D):
            self.df.data4D[k] = take(self.df.data4D[k], 2)
        for k in list(self.df.dataFullSpec):
            self.df.dataFullSpec[k] = take(self.df.dataFullSpec[k], 3)

    def filterProfiles(self, condition):
        """Keeps the columns where the 2D boolean `condition` holds; the grid
        becomes (n, 1) (core.py:1001-1078)."""
        condition = np.asarray(condition)
        if condition.shape != self._shape2D and condition.shape != (int(np.prod(self._shape2D)),):
            raise ValueError("shape mismatch, shape of condition must be 2D field!")
        condition = condition.reshape(self._shape2D)
        assert np.sum(condition) >= 1
        n = int(np.sum(condition))
        self._reshape_all(lambda a, t: a[condition].reshape((n, 1) + a.shape[2:]))
        self.p["ngridx"], self.p["ngridy"] = n, 1
        self._set_shapes()

    def tileProfiles(self, rep2D):
        """Repeats the columns rep2D = (rx, ry) times (core.py:1080-1133)."""
        assert len(rep2D) == 2
        self._reshape_all(lambda a, t: np.tile(a, tuple(rep2D) + (1,) * t))
        self.p["ngridx"] *= rep2D[0]
        self.p["ngridy"] *= rep2D[1]
        self._set_shapes()

    def addProfile(self, profile, axis=0):
        """Appends the columns of another profile dict along `axis`; the
        layer counts must agree (core.py:1135-1159)."""
        for k in list(self.p):
            if np.ndim(self.p[k]) >= 2:
                self.p[k] = np.concatenate((self.p[k], profile[k]), axis=axis)
        self.p["ngridx"], self.p["ngridy"] = np.shape(self.p["hgt_lev"])[:2]
        self._set_shapes()

    def writeNmlFile(self, nmlFile):
        """Writes nmlSet as a Fortran namelist "&settings ... /"
        (core.py:277-305)."""
        with open(nmlFile, "w") as f:
            f.write("&settings\n\r")
            for key, v in self.nmlSet.items():
                if isinstance(v, (bool, np.bool_)):
                    f.write("%s=.%s.\n\r" % (key, str(bool(v)).lower()))
                elif isinstance(v, (int, np.integer)):
                    f.write("%s=%i\n\r" % (key, int(v)))
                elif isinstance(v, (float, np.floating)):
                    f.write("%s=%f\n\r" % (key, np.float64(v)))
                elif isinstance(v, str):
                    f.write("%s='%s'\n\r" % (key, v))
                else:
                    raise ValueError("cannot determine type of nml key " + key)
            f.write("/\n\r")

    def readNmlFile(self, inputFile):
        """Reads a Fortran namelist file into nmlSet: numbers (d exponents
        too), .true./.false., strings; keys not among the defaults are
        skipped (core.py:307-335)."""
        import os
        import re
        if not os.path.exists(inputFile):
            raise IOError("file not found: " + inputFile)
        text = open(inputFile).read()
        defaults = pyPamtra().nmlSet
        # a group ends at a "/" that starts a line (values may hold paths)
        for group in re.findall(r"&\w+(.*?)[\n\r]\s*/", text, flags=re.S):
            for key, raw in re.findall(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^,\s]+)", group):
                key = key.lower()
                if key not in defaults:
                    if self.set["pyVerbose"] > 0:
                        print("Setting '" + key + "' from '" + inputFile + "' skipped.")
                    continue
                raw = raw.strip()
                if raw.lower() in (".true.", "t", ".t."):
                    value = True
                elif raw.lower() in (".false.", "f", ".f."):
                    value = False
                elif raw[0] in "'\"":
                    value = raw[1:-1]
                else:
                    try:
                        value = float(raw.replace("d", "e").replace("D", "e"))
                    except ValueError:
                        value = raw
                self.nmlSet[key] = value

    def _calcMoistRho(self):
        from . import meteoSI
        h = dict()
        h["dz"] = self.p["hgt_lev"][..., 1:] - self.p["hgt_lev"][..., :-1]
        rh = self.p["relhum"] / 100.
        h["vapor"] = meteoSI.rh2q(rh, self.p["temp"], self.p["press"])
        h["sum_hydro_q"] = np.nansum(self.p["hydro_q"], axis=-1)
        h["rho_moist"] = meteoSI.moist_rho_rh(self.p["press"], self.p["temp"], rh,
                                              h["sum_hydro_q"])
        self._helperP = h

    def addIntegratedValues(self):
        """iwv and the hydrometeor paths hydro_wp [kg/m2] (core.py:1337-1349);
        needs the layer fields temp, press, relhum [%]."""
        self._shape3Dhyd = (self.p["ngridx"], self.p["ngridy"], self.df.nhydro)
        self._calcMoistRho()
        h = self._helperP
        self.p["iwv"] = np.nansum(h["vapor"] * h["rho_moist"] * h["dz"], axis=-1)
        self.p["hydro_wp"] = np.zeros(self._shape3Dhyd)
        if not np.all(self.p["hydro_q"] == 0):
            for i in range(self.df.nhydro):
                self.p["hydro_wp"][..., i] = np.nansum(
                    self.p["hydro_q"][..., i] * h["rho_moist"] * h["dz"], axis=-1)

    def runParallelPamtra(self, freqs, pp_local_workers="auto", pp_deltaF=1, pp_deltaX=0,
                          pp_deltaY=0, checkData=True, timeout=None):
        """The reference splits the run over local worker processes
        (core.py:1570-1675); here the device already runs every column and
        channel of one call in parallel, so this is runPamtra with the same
        results."""
        return self.runPamtra(freqs, checkData=checkData)

    def averageResultTbs(self, translatorDict):
        """Averages the TB of channels given as several frequencies, e.g.
        {120.15: [117.35, 120.15]}; keeps the originals in
        r["not_averaged_tb"] and set["not_averaged_freqs"] (core.py:2555-2593)."""
        assert not self.nmlSet["active"]
        assert self.nmlSet["passive"]
        self.r["not_averaged_tb"] = deepcopy(self.r["tb"])
        self.set["not_averaged_freqs"] = deepcopy(self.set["freqs"])
        self.set["freqs"] = sorted(translatorDict.keys())
        self.set["nfreqs"] = len(self.set["freqs"])
        tb0 = self.r["not_averaged_tb"]
        self.r["tb"] = np.ones(tb0.shape[:4] + (self.set["nfreqs"], tb0.shape[5])) * missingNumber
        old = np.array(self.set["not_averaged_freqs"])
        for ff, (freqNew, freqList) in enumerate(sorted(translatorDict.items())):
            idx = []
            for fq in freqList:
                found = np.where(fq == old)[0]
                assert len(found) == 1
                idx.append(found[0])
            self.r["tb"][:, :, :, :, ff, :] = np.mean(tb0[:, :, :, :, idx, :], axis=4)

    def writeResultsToNumpy(self, fname, seperateFiles=False):
        """Saves results, profile, settings and descriptors: one pickle
        file, or one .npy file per entry in the directory `fname`
        (core.py:2072-2114)."""
        import os
        import pickle
        if not seperateFiles:
            with open(fname, "wb") as f:
                pickle.dump([self.r, self.p, self.nmlSet, self.set, self.df.data, self.df.data4D,
                             self.df.dataFullSpec], f)
            return
        os.makedirs(fname)
        for name, dic in (("r", self.r), ("p", self.p), ("nmlSet", self.nmlSet), ("set", self.set),
                          ("df.data4D", self.df.data4D), ("df.dataFullSpec", self.df.dataFullSpec)):
            for key, data in dic.items():
                path = fname + "/" + name + "%" + key + "%" + ".npy"
                if isinstance(data, np.ma.MaskedArray):
                    data = data.filled(-9999)
                if isinstance(data, np.ndarray) and data.dtype == np.float64:
                    data = data.astype("f4")
                elif isinstance(data, np.ndarray) and data.dtype == np.int64:
                    data = data.astype("i4")
                np.save(path, np.asarray(data) if not isinstance(data, (dict, list)) else
                        np.array(data, dtype=object), allow_pickle=True)

    def loadResultsFromNumpy(self, fname):
        """Loads what writeResultsToNumpy wrote (core.py:2116-2161).  Only
        for files this library wrote: they are pickles."""
        import os
        import pickle
        if os.path.isdir(fname):
            self.r, self.p, self.set = dict(), dict(), dict()
            self.df.data4D, self.df.dataFullSpec = dict(), dict()
            for fn in os.listdir(fname):
                key, sub, _ = fn.split("%")
                v = np.load(fname + "/" + fn, allow_pickle=True)
                v = v.item() if v.dtype == object and v.ndim == 0 else v
                target = {"r": self.r, "p": self.p, "nmlSet": self.nmlSet, "set": self.set,
                          "df.data4D": self.df.data4D,
                          "df.dataFullSpec": self.df.dataFullSpec}[key]
                target[sub] = v[()] if isinstance(v, np.ndarray) and v.ndim == 0 else v
            return
        with open(fname, "rb") as f:
            (self.r, self.p, self.nmlSet, self.set, self.df.data, self.df.data4D,
             self.df.dataFullSpec) = pickle.load(f)
        self.df.nhydro = len(self.df.data)
        self.p.setdefault("noutlevels", 2)
        self._set_shapes()

    def addCloudShape(self):
        """cloudBase and cloudTop [m] of the liquid clouds that
        meteoSI.detect_liq_cloud finds on the first nlyrs levels; with several
        layers the lowest base and the uppermost top (core.py:1360-1375)."""
        from .meteoSI import detect_liq_cloud
        self.p["cloudTop"] = np.ones(self._shape2D) * missingNumber
        self.p["cloudBase"] = np.ones(self._shape2D) * missingNumber
        for x in range(self._shape2D[0]):
            for y in range(self._shape2D[1]):
                n = int(self.p["nlyrs"][x, y])
                i_top, i_base, _ = detect_liq_cloud(self.p["hgt_lev"][x, y, :n],
                                                    self.p["temp_lev"][x, y, :n],
                                                    self.p["relhum_lev"][x, y, :n])
                if len(i_top) > 0:
                    self.p["cloudTop"][x, y] = self.p["hgt_lev"][x, y, i_top[-1] + 1]
                    self.p["cloudBase"][x, y] = self.p["hgt_lev"][x, y, i_base[0]]

    def createFullProfile(self, timestamp, lat, lon, wind10u, wind10v, obs_height, hgt_lev,
                          press_lev, temp_lev, relhum_lev, hydro_q, hydro_n, hydro_reff, radar_prop,
                          sfc_type, sfc_model, sfc_refl, sfc_salinity, sfc_slf, sfc_sif):
        """A complete profile from arrays of one, two or three grid
        dimensions, only reshaped: nothing is derived or filled
        (core.py:1449-1516).  model_i/model_j are the 1-based grid indices."""
        nd = np.ndim(hgt_lev)
        if nd == 1:
            nx, ny = 1, 1
        elif nd == 2:
            nx, ny = np.shape(hgt_lev)[0], 1
        elif nd == 3:
            nx, ny = np.shape(hgt_lev)[:2]
        else:
            raise IOError("Too many dimensions!")
        p = self.p
        p["ngridx"], p["ngridy"] = nx, ny
        p["nlyrs"] = (np.sum(np.asarray(hgt_lev) != missingNumber, axis=-1) - 1).reshape(nx, ny)
        p["max_nlyrs"] = np.shape(hgt_lev)[-1] - 1
        p["noutlevels"] = np.shape(obs_height)[-1]
        self._set_shapes()
        s2, s3o, s3p, s4 = self._shape2D, self._shape3Dout, self._shape3Dplus, self._shape4D
        for k, v, shp in (("unixtime", timestamp, s2), ("lat", lat, s2), ("lon", lon, s2),
                          ("wind10u", wind10u, s2), ("wind10v", wind10v, s2),
                          ("obs_height", obs_height, s3o), ("hydro_q", hydro_q, s4),
                          ("hydro_n", hydro_n, s4), ("hydro_reff", hydro_reff, s4),
                          ("hgt_lev", hgt_lev, s3p), ("temp_lev", temp_lev, s3p),
                          ("press_lev", press_lev, s3p), ("relhum_lev", relhum_lev, s3p),
                          ("radar_prop", radar_prop, s2), ("sfc_type", sfc_type, s2),
                          ("sfc_model", sfc_model, s2), ("sfc_refl", sfc_refl, s2),
                          ("sfc_salinity", sfc_salinity, s2), ("sfc_slf", sfc_slf, s2),
                          ("sfc_sif", sfc_sif, s2)):
            p[k] = np.asarray(v).reshape(shp)
        gi, gj = np.indices(s2)
        p["model_i"], p["model_j"] = gi + 1, gj + 1
