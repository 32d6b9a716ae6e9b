"""ctypes mirror of include/pamtra_amd.h and the marshalling of pyPamtra
dictionaries into it.

The reference pushes the same information into Fortran module globals through
f2py (python/pyPamtra/libWrapper.py:24-235); here it becomes flat C-order
arrays plus a few structs.  Arrays may be numpy arrays (host, for ``pam_run``)
or torch tensors already on the device (for ``pam_run_device``).
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

NUMMU = 16  # the reference's compile-time nummu (settings.f90:16)
NSTOKES = 2
NANGLES = 2 * NUMMU
# angles per hemisphere -> the library built for it (pyPamtra's _nangles)
LIBRARIES = {16: "libpamtra_amd.so", 32: "libpamtra_amd_n64.so"}

PAM_SUCCESS, PAM_WARNING, PAM_FATAL, PAM_INFO = 0, 1, 2, 3
NOALLOC_SCRATCH_BYTES = 256 << 20  # PAM_NOALLOC_SCRATCH_BYTES

DIST_NAMES = {
    "disabled": 0,
    "mono": 1,
    "exp": 2,
    "exp_field_t": 3,
    "exp_ryan": 4,
    "mgamma": 5,
    "logn": 6,
    "exp_cosmo_snow": 7,
    "mono_cosmo_ice": 8,
    "mono_echam_cl": 9,
    "mono_echam_ice": 10,
    "mgamma_MNH": 11,
    "norm_gamma": 12,
    "norm_mgamma": 13,
}
# distributions the reference names but whose bins it never fills:
# make_dist.f90:164-215 sets f_ds for neither, so its result is undefined
UNDEFINED_DISTS = ("const", "const_cosmo_ice")
SCAT_NAMES = {"disabled": 0, "mie-sphere": 1, "ssrg-rt3": 2, "tmatrix": 3}
PRECISIONS = {"fp64": 0, "mixed": 1}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class PamSettings(C.Structure):
    _fields_ = [
        ("passive", C.c_int32),
        ("active", C.c_int32),
        ("lgas_extinction", C.c_int32),
        ("lhyd_absorption", C.c_int32),
        ("lhyd_scattering", C.c_int32),
        ("lhyd_emission", C.c_int32),
        ("hydro_includehydroinrhoair", C.c_int32),
        ("hydro_limit_density_area", C.c_int32),
        ("hydro_adaptive_grid", C.c_int32),
        ("conserve_mass_rescale_dsd", C.c_int32),
        ("gas_mod", C.c_int32),
        ("liq_mod", C.c_int32),
        ("outpol", C.c_int32),
        ("precision", C.c_int32),
        ("hydro_threshold", C.c_double),
        ("hydro_softsphere_min_density", C.c_double),
        ("emissivity", C.c_double),
        ("radar_attenuation", C.c_int32),
        ("radar_npol", C.c_int32),
        ("radar_pol", C.c_int32 * 5),
        ("radar_pad_", C.c_int32),
        ("radar_k2", _dp),
    ]


RADAR_ATTENUATION = {"disabled": 0, "top-down": 1, "bottom-up": 2}
RADAR_POL = {"NN": 0, "HH": 1, "VV": 2, "HV": 3, "VH": 4}


class PamHydroDesc(C.Structure):
    _fields_ = [
        ("as_ratio", C.c_double),
        ("liq_ice", C.c_int32),
        ("moment_in", C.c_int32),
        ("rho_ms", C.c_double),
        ("a_ms", C.c_double),
        ("b_ms", C.c_double),
        ("alpha_as", C.c_double),
        ("beta_as", C.c_double),
        ("nbin", C.c_int32),
        ("dist_name", C.c_int32),
        ("p_1", C.c_double),
        ("p_2", C.c_double),
        ("p_3", C.c_double),
        ("p_4", C.c_double),
        ("d_1", C.c_double),
        ("d_2", C.c_double),
        ("scat_name", C.c_int32),
        ("pad_", C.c_int32),
        ("canting", C.c_double),
        ("rg_kappa", C.c_double),
        ("rg_beta", C.c_double),
        ("rg_gamma", C.c_double),
        ("rg_zeta", C.c_double),
    ]


HYDRO4D_KEYS = ["as_ratio", "rho_ms", "a_ms", "b_ms", "alpha_as", "beta_as",
                "p_1", "p_2", "p_3", "p_4", "d_1", "d_2"]


# full-spectrum arrays (descriptorFile.addFullSpectra) -> pam_hydro_4d fields;
# area_ds and fallvelocity enter only the radar spectra, which are not built
FULLSPEC_KEYS = {"rho_ds": "fs_rho", "d_ds": "fs_d", "d_bound_ds": "fs_d_bound", "n_ds": "fs_n",
                 "mass_ds": "fs_mass", "as_ratio": "fs_as_ratio", "canting": "fs_canting",
                 "rg_kappa_ds": "fs_rg_kappa", "rg_beta_ds": "fs_rg_beta",
                 "rg_gamma_ds": "fs_rg_gamma", "rg_zeta_ds": "fs_rg_zeta"}


class PamHydro4D(C.Structure):
    _fields_ = [(k, _dp) for k in HYDRO4D_KEYS] + [("nbin", _ip), ("fs_nbin", C.c_int32),
                                                   ("fs_pad_", C.c_int32)] + \
        [(v, _dp) for v in FULLSPEC_KEYS.values()]


PROFILE_DOUBLE_KEYS = ["hgt_lev", "temp_lev", "press_lev", "relhum_lev", "hgt", "temp",
                       "press", "relhum", "hydro_q", "hydro_n", "hydro_reff", "groundtemp",
                       "wind10u", "wind10v", "sfc_salinity", "sfc_emissivity"]


class PamProfile(C.Structure):
    _fields_ = [
        ("ngridx", C.c_int32),
        ("ngridy", C.c_int32),
        ("max_nlyrs", C.c_int32),
        ("noutlevels", C.c_int32),
        ("nhydro", C.c_int32),
        ("pad_", C.c_int32),
        ("nlyrs", _ip),
    ] + [(k, _dp) for k in PROFILE_DOUBLE_KEYS[:16]] + [
        ("sfc_type", _ip),
        ("sfc_model", _ip),
        ("sfc_refl", C.c_char_p),
        ("obs_height", _dp),
    ]


class PamResults(C.Structure):
    _fields_ = [
        ("tb", _dp),
        ("emissivity", _dp),
        ("angles_deg", _dp),
        ("status", _ip),
        ("kextatmo", _dp),
        ("kexttot", _dp),
        ("salbedo", _dp),
        ("nlegen", _ip),
        ("ndoubles", _ip),
        ("msg", C.c_char_p),
        ("msglen", C.c_int32),
        ("pad_", C.c_int32),
        ("ze", _dp),
        ("att_hydro", _dp),
        ("att_atmo", _dp),
        ("radar_hgt", _dp),
    ]


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a, ctype):
    """Pointer to the data of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    if _is_torch(a):
        return C.cast(C.c_void_p(a.data_ptr()), C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


def _check_tensor(a, dtype_name, shape):
    """A device tensor handed to pam_run_device must already be what the
    kernels read: the exact dtype, contiguous, on the GPU and of the expected
    element count.  Anything else would be read as raw memory of that type."""
    if str(a.dtype) != "torch." + dtype_name:
        raise ValueError(f"tensor of dtype {a.dtype}: {dtype_name} is required")
    if not a.is_contiguous():
        raise ValueError("tensor is not contiguous")
    if shape is not None and a.numel() != int(np.prod(shape)):
        raise ValueError(f"tensor of {a.numel()} elements: shape {tuple(shape)} is expected")
    if a.device.type != "cuda":
        raise ValueError(f"tensor on {a.device}: a device tensor is required")
    return a


def _as_f64(a, shape=None):
    if a is None:
        return None
    if _is_torch(a):
        return _check_tensor(a, "float64", shape)
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        arr = arr.reshape(shape)
    return arr


def _as_i32(a, shape=None):
    if a is None:
        return None
    if _is_torch(a):
        return _check_tensor(a, "int32", shape)
    arr = np.ascontiguousarray(np.asarray(a).astype(np.int32))
    if shape is not None:
        arr = arr.reshape(shape)
    return arr


def _flag(v) -> int:
    return 1 if bool(v) else 0


def make_settings(nmlSets: dict, sets: dict | None = None) -> PamSettings:
    """nmlSet -> pam_settings.  Options outside the passive Mie path fail
    loudly, as the reference's Fortran would for an unknown model name."""
    s = PamSettings()
    g = nmlSets.get
    s.passive = _flag(g("passive", True))
    s.active = _flag(g("active", True))
    s.lgas_extinction = _flag(g("lgas_extinction", True))
    s.lhyd_absorption = _flag(g("lhyd_absorption", True))
    s.lhyd_scattering = _flag(g("lhyd_scattering", True))
    s.lhyd_emission = _flag(g("lhyd_emission", True))
    s.hydro_includehydroinrhoair = _flag(g("hydro_includehydroinrhoair", True))
    s.hydro_limit_density_area = _flag(g("hydro_limit_density_area", True))
    s.hydro_adaptive_grid = _flag(g("hydro_adaptive_grid", True))
    s.conserve_mass_rescale_dsd = _flag(g("conserve_mass_rescale_dsd", True))
    gas = str(g("gas_mod", "R98")).strip()
    if gas != "R98":
        raise NotImplementedError(f"gas_mod {gas!r}: only 'R98' is implemented")
    s.gas_mod = 0
    liq = str(g("liq_mod", "TKC")).strip()
    if liq != "TKC":
        raise NotImplementedError(f"liq_mod {liq!r}: only 'TKC' is implemented")
    s.liq_mod = 0
    outpol = str(g("outpol", "VH")).strip()
    if outpol not in ("VH", "IQ"):
        raise ValueError(f"outpol {outpol!r} must be 'VH' or 'IQ'")
    s.outpol = 0 if outpol == "VH" else 1
    prec = (sets or {}).get("precision", "fp64")
    if prec not in PRECISIONS:
        raise ValueError(f"precision {prec!r} must be one of {sorted(PRECISIONS)}")
    s.precision = PRECISIONS[prec]
    s.hydro_threshold = float(g("hydro_threshold", 1e-10))
    s.hydro_softsphere_min_density = float(g("hydro_softsphere_min_density", 10.0))
    s.emissivity = float(g("emissivity", 0.6))
    # radar simulator (src/radar_simulator.f90): the "simple" mode only
    if s.active:
        mode = str(g("radar_mode", "simple")).strip()
        if mode != "simple":
            raise NotImplementedError(f"radar_mode {mode!r}: only 'simple' is implemented")
        if int(g("radar_npeaks", 1)) != 1:
            raise ValueError("radar_mode 'simple' needs radar_npeaks=1 "
                             "(radar_simulator.f90:161-163)")
    att = str(g("radar_attenuation", "disabled")).strip()
    if att not in RADAR_ATTENUATION:
        raise ValueError(f"radar_attenuation {att!r} must be one of {sorted(RADAR_ATTENUATION)}")
    s.radar_attenuation = RADAR_ATTENUATION[att]
    pols = [x.strip() for x in str(g("radar_polarisation", "NN")).split(",") if x.strip()]
    if not 1 <= len(pols) <= 5 or any(x not in RADAR_POL for x in pols):
        raise ValueError(f"radar_polarisation {g('radar_polarisation')!r}: 1-5 of "
                         f"{sorted(RADAR_POL)}")
    s.radar_npol = len(pols)
    for i, x in enumerate(pols):
        s.radar_pol[i] = RADAR_POL[x]
    return s


def make_hydros(descriptorFile) -> C.Array:
    """descriptor recarray (descriptorFile.py:19-20 layout) -> pam_hydro_desc[]."""
    n = len(descriptorFile)
    arr = (PamHydroDesc * max(n, 1))()
    names = descriptorFile.dtype.names
    for i in range(n):
        row = descriptorFile[i]
        h = arr[i]

        def num(key, default=-99.0):
            return float(row[key]) if key in names else default

        dist = str(row["dist_name"]).strip().strip("'\"")
        if dist in UNDEFINED_DISTS:
            raise NotImplementedError(
                f"dist_name {dist!r}: the reference's make_dist leaves its bins unset "
                "(make_dist.f90:164-215), so it has no defined result to reproduce")
        if dist not in DIST_NAMES:
            raise NotImplementedError(f"dist_name {dist!r} is not implemented")
        if dist in ("mono_echam_cl", "mono_echam_ice") and "nbin" in names and int(row["nbin"]) != 2:
            # hydrometeor_extinction.f90:200-202 forces nbin = 2 only for mono and
            # mono_cosmo_ice; make_dist.f90:89-97 fills three edges either way
            raise ValueError(f"dist_name {dist!r} needs nbin = 2 (the reference reads three "
                             "bin edges only)")
        scat = str(row["scat_name"]).strip()
        scat, rg = scattering_model(scat)
        h.as_ratio = num("as_ratio")
        h.liq_ice = int(row["liq_ice"])
        h.moment_in = int(row["moment_in"])
        h.rho_ms = num("rho_ms")
        h.a_ms = num("a_ms")
        h.b_ms = num("b_ms")
        h.alpha_as = num("alpha_as")
        h.beta_as = num("beta_as")
        h.nbin = int(row["nbin"]) if "nbin" in names else 0
        h.dist_name = DIST_NAMES[dist]
        h.p_1, h.p_2, h.p_3, h.p_4 = num("p_1"), num("p_2"), num("p_3"), num("p_4")
        h.d_1, h.d_2 = num("d_1"), num("d_2")
        h.scat_name = SCAT_NAMES[scat]
        h.canting = num("canting")
        h.rg_kappa, h.rg_beta, h.rg_gamma, h.rg_zeta = rg
    return arr


# self-similar Rayleigh-Gans defaults for aggregates (Hogan and Westbrook
# 2014), zeta as the single-precision literal 0.04 (scatProperties.f90:793-797)
SSRG_DEFAULTS = (0.25, 1.0, 1.66, float(np.float32(0.04)))


def scattering_model(name: str):
    """scat_name -> (model key of SCAT_NAMES, (kappa, beta, gamma, zeta)).

    "ssrg-rt3" takes its parameters from the name as
    ssrg-rt3_<kappa>_<beta>_<gamma>[_<zeta>], zeta 1.0 when left out
    (scatProperties.f90:771-798).  The reference reads only the first eight
    characters for the model and the rest from the tenth on; a name with
    fewer than three or more than four parameters leaves it reading an unset
    token, which is refused here."""
    if name in ("disabled", "mie-sphere", "tmatrix"):
        return name, (0.0, 0.0, 0.0, 0.0)
    if name[:8] == "ssrg-rt3":
        if len(name) == 8:
            return "ssrg-rt3", SSRG_DEFAULTS
        tok = name[9:].split("_")
        if len(tok) not in (3, 4):
            raise ValueError(f"scat_name {name!r}: ssrg-rt3 takes kappa_beta_gamma[_zeta]")
        try:
            vals = [float(t) for t in tok]
        except ValueError:
            raise ValueError(f"scat_name {name!r}: parameters are not numbers") from None
        if len(vals) == 3:
            vals.append(1.0)
        return "ssrg-rt3", tuple(vals)
    if name[:5] == "ssrga":
        why = "the reference's calc_single_ssrga leaves its amplitude matrix unset"
    elif name[:16] == "ss-rayleigh-gans":
        why = "the reference's passive routine asserts it is unimplemented (rayleigh_gans.f90:1019-1020)"
    elif name in ("rayleigh", "rayleigh-gans"):
        why = ("a backscatter-only model with zero extinction and albedo in the reference "
               "(scatProperties.f90:625-720); not built")
    elif name[:5] in ("liudb", "hongd"):
        why = "a DDA-database model that needs the reference's external data files; not built"
    else:
        why = "unknown scattering model"
    raise NotImplementedError(f"scat_name {name!r}: {why}; 'mie-sphere', 'ssrg-rt3' and "
                              "'tmatrix' are implemented")


def check_tmatrix(hydros, settings, nmlSets):
    """The T-matrix producer's limits (scatProperties.f90:289-368).  Mixing it
    with the Legendre-series models is refused because the reference's result
    is then undefined: the T-matrix branch adds salb_hydro * kext_hydro to the
    layer albedo without ever setting salb_hydro (scatProperties.f90:141,
    911).  The radar simulator is not implemented with it on the device.  tmatrix_db 'file' (the
    reference's on-disk cache of the same results) computes them."""
    names = [int(h.scat_name) for h in hydros]
    if SCAT_NAMES["tmatrix"] not in names:
        return
    if any(n not in (SCAT_NAMES["tmatrix"], SCAT_NAMES["disabled"]) for n in names):
        raise NotImplementedError(
            "scat_name 'tmatrix' mixed with 'mie-sphere' or 'ssrg-rt3': the reference's layer "
            "albedo then reads an unset value (scatProperties.f90:141,911), so that mix has no "
            "defined result")
    if settings.active:
        raise NotImplementedError("scat_name 'tmatrix' with active (radar) simulation is not "
                                  "implemented")
    db = str(nmlSets.get("tmatrix_db", "none")).strip()
    if db not in ("none", "file"):
        raise ValueError(f"tmatrix_db {db!r} must be 'none' or 'file'")


@dataclass
class Marshalled:
    settings: PamSettings
    hydros: object
    hydro4d: PamHydro4D
    profile: PamProfile
    nx: int
    ny: int
    max_nlyrs: int
    noutlevels: int
    nhydro: int
    keep: list = field(default_factory=list)
    radar_k2: object = 0.93  # nmlSet radar_k2: one value or one per frequency
    nummu: int = NUMMU       # quadrature angles per hemisphere (sets["_nangles"])


def marshal(nmlSets, sets, descriptorFile, descriptorFile4D, profile,
            descriptorFileFS=None) -> Marshalled:
    """Build the C structs from the pyPamtra dictionaries.  The returned
    object keeps every array alive for the duration of the call.  With
    nmlSet hydro_fullspec the full-spectrum arrays of descriptorFileFS
    (descriptorFile.addFullSpectra) replace the size distributions."""
    keep = []
    nx, ny = int(profile["ngridx"]), int(profile["ngridy"])
    L = int(profile["max_nlyrs"])
    nout = int(profile.get("noutlevels", 2))
    nh = len(descriptorFile)
    settings = make_settings(nmlSets, sets)
    hydros = make_hydros(descriptorFile)
    keep.append(hydros)
    check_tmatrix(hydros, settings, nmlSets)
    h4 = PamHydro4D()
    for key, arr in (descriptorFile4D or {}).items():
        if key == "nbin":
            a = _as_i32(arr, (nx, ny, L, nh))
            keep.append(a)
            h4.nbin = _ptr(a, C.c_int32)
        elif key in HYDRO4D_KEYS:
            a = _as_f64(arr, (nx, ny, L, nh))
            keep.append(a)
            setattr(h4, key, _ptr(a, C.c_double))
        else:
            raise NotImplementedError(f"4D descriptor field {key!r} is not supported")
    if nmlSets.get("hydro_fullspec", False):
        fs = descriptorFileFS or {}
        missing = [k for k in FULLSPEC_KEYS if k not in fs]
        if missing:
            raise ValueError(f"hydro_fullspec needs df.addFullSpectra(); missing {missing}")
        nb = int(np.shape(fs["d_ds"])[-1])
        if nb < 1:
            raise ValueError("hydro_fullspec: no bins")
        h4.fs_nbin = nb
        for key, fld in FULLSPEC_KEYS.items():
            a = _as_f64(fs[key], (nx, ny, L, nh, nb + 1 if key == "d_bound_ds" else nb))
            keep.append(a)
            setattr(h4, fld, _ptr(a, C.c_double))
    p = PamProfile()
    p.ngridx, p.ngridy, p.max_nlyrs, p.noutlevels, p.nhydro = nx, ny, L, nout, nh
    shapes = {
        "hgt_lev": (nx, ny, L + 1), "temp_lev": (nx, ny, L + 1), "press_lev": (nx, ny, L + 1),
        "relhum_lev": (nx, ny, L + 1), "hgt": (nx, ny, L), "temp": (nx, ny, L),
        "press": (nx, ny, L), "relhum": (nx, ny, L), "hydro_q": (nx, ny, L, nh),
        "hydro_n": (nx, ny, L, nh), "hydro_reff": (nx, ny, L, nh), "groundtemp": (nx, ny),
        "wind10u": (nx, ny), "wind10v": (nx, ny), "sfc_salinity": (nx, ny),
        "sfc_emissivity": (nx, ny),
    }
    for key in PROFILE_DOUBLE_KEYS:
        v = profile.get(key)
        if v is None or (key.startswith("hydro_") and nh == 0):
            continue
        a = _as_f64(v, shapes[key])
        keep.append(a)
        setattr(p, key, _ptr(a, C.c_double))
    nl = _as_i32(profile["nlyrs"], (nx, ny))
    keep.append(nl)
    p.nlyrs = _ptr(nl, C.c_int32)
    for key in ("sfc_type", "sfc_model"):
        v = profile.get(key)
        if v is not None:
            a = _as_i32(v, (nx, ny))
            keep.append(a)
            setattr(p, key, _ptr(a, C.c_int32))
    refl = profile.get("sfc_refl")
    if refl is not None:
        if _is_torch(refl):
            _check_tensor(refl, "uint8", (nx, ny))
            keep.append(refl)
            p.sfc_refl = C.cast(C.c_void_p(refl.data_ptr()), C.c_char_p)
        else:
            r = np.asarray(refl)
            if r.dtype.kind in ("U", "S"):
                flat = [str(x.decode() if isinstance(x, bytes) else x)[:1] or "S"
                        for x in r.reshape(-1)]
                buf = C.create_string_buffer("".join(flat).encode("ascii"), nx * ny + 1)
            else:
                buf = C.create_string_buffer(r.astype(np.uint8).tobytes(), nx * ny + 1)
            keep.append(buf)
            p.sfc_refl = C.cast(buf, C.c_char_p)
    oh = profile.get("obs_height")
    if oh is not None:
        a = _as_f64(oh, (nx, ny, nout))
        keep.append(a)
        p.obs_height = _ptr(a, C.c_double)
    return Marshalled(settings, hydros, h4, p, nx, ny, L, nout, nh, keep,
                      (nmlSets or {}).get("radar_k2", 0.93), nangles(sets))


@dataclass
class HostResults:
    struct: PamResults
    arrays: dict
    keep: list


def alloc_results(m: Marshalled, nfreq: int, diagnostics: bool = False,
                  msglen: int = 512) -> HostResults:
    nx, ny, L, nout = m.nx, m.ny, m.max_nlyrs, m.noutlevels
    nmu = m.nummu
    arrs = {
        "tb": np.full((nx, ny, nout, 2 * nmu, nfreq, NSTOKES), np.nan),
        "emissivity": np.full((nx, ny, NSTOKES, nfreq, nmu), np.nan),
        "angles_deg": np.full((2 * nmu,), np.nan),
        "status": np.zeros((nx, ny, nfreq), dtype=np.int32),
    }
    if m.settings.active:
        npol = int(m.settings.radar_npol)
        arrs["ze"] = np.full((nx, ny, L, nfreq, npol, 1), -9999.0)
        arrs["att_hydro"] = np.zeros((nx, ny, L, nfreq, 1))
        arrs["att_atmo"] = np.zeros((nx, ny, L, nfreq))
        arrs["radar_hgt"] = np.full((nx, ny, L), -9999.0)
        k2 = np.asarray(m.radar_k2, dtype=np.float64).reshape(-1)
        if k2.size == 1:
            k2 = np.full(nfreq, float(k2[0]))
        if k2.size < nfreq:
            raise ValueError(f"radar_k2 has {k2.size} values for {nfreq} frequencies")
        k2 = np.ascontiguousarray(k2[:nfreq])
        m.keep.append(k2)
        m.settings.radar_k2 = k2.ctypes.data_as(C.POINTER(C.c_double))
    if diagnostics:
        arrs["kextatmo"] = np.full((nx, ny, L, nfreq), np.nan)
        arrs["kexttot"] = np.full((nx, ny, L, nfreq), np.nan)
        arrs["salbedo"] = np.full((nx, ny, L, nfreq), np.nan)
        arrs["nlegen"] = np.full((nx, ny, L, nfreq), -9999, dtype=np.int32)
        arrs["ndoubles"] = np.full((nx, ny, L, nfreq), -9999, dtype=np.int32)
    r = PamResults()
    for k, a in arrs.items():
        setattr(r, k, _ptr(a, C.c_int32 if a.dtype == np.int32 else C.c_double))
    msg = C.create_string_buffer(msglen)
    r.msg = C.cast(msg, C.c_char_p)
    r.msglen = msglen
    return HostResults(r, arrs, [msg])


def results_message(hr: HostResults) -> str:
    return hr.keep[0].value.decode(errors="replace")


def freqs_array(freqs):
    f = np.ascontiguousarray(np.asarray(freqs, dtype=np.float64).reshape(-1))
    return f, f.ctypes.data_as(C.POINTER(C.c_double))


def declare_library(lib: C.CDLL) -> C.CDLL:
    """Argument and return types of the product library's C-ABI."""
    lib.pam_version.restype = C.c_char_p
    lib.pam_version.argtypes = []
    lib.pam_nummu.restype = C.c_int32
    lib.pam_nummu.argtypes = []
    lib.pam_device_count.restype = C.c_int
    lib.pam_device_count.argtypes = []
    lib.pam_context_create.restype = C.c_int
    lib.pam_context_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
    lib.pam_context_destroy.restype = C.c_int
    lib.pam_context_destroy.argtypes = [C.c_void_p]
    run_args = [C.c_void_p, C.POINTER(PamSettings), C.POINTER(PamHydroDesc),
                C.POINTER(PamHydro4D), C.POINTER(PamProfile), C.POINTER(C.c_double),
                C.c_int32, C.POINTER(PamResults)]
    lib.pam_run.restype = C.c_int
    lib.pam_run.argtypes = run_args
    lib.pam_run_device.restype = C.c_int
    lib.pam_run_device.argtypes = run_args + [C.c_void_p]
    lib.pam_run_device_noalloc.restype = C.c_int
    lib.pam_run_device_noalloc.argtypes = run_args + [C.c_void_p]
    lib.pam_last_stage_times.restype = C.c_int
    lib.pam_last_stage_times.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int32]
    lib.pam_stage_names.restype = C.c_char_p
    lib.pam_stage_names.argtypes = []
    lib.pam_last_work.restype = C.c_int
    lib.pam_last_work.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32]
    lib.pam_workspace_bytes.restype = C.c_int64
    lib.pam_workspace_bytes.argtypes = [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_int64]
    lib.pam_context_set_workspace_limit.restype = C.c_int
    lib.pam_context_set_workspace_limit.argtypes = [C.c_void_p, C.c_int64]
    lib.pam_context_bytes.restype = C.c_int64
    lib.pam_context_bytes.argtypes = [C.c_void_p]
    lib.pam_context_trim.restype = C.c_int
    lib.pam_context_trim.argtypes = [C.c_void_p]
    lib.pam_tuning_keys.restype = C.c_char_p
    lib.pam_tuning_keys.argtypes = []
    lib.pam_context_set_tuning.restype = C.c_int
    lib.pam_context_set_tuning.argtypes = [C.c_void_p, C.c_char_p, C.c_int64]
    lib.pam_context_get_tuning.restype = C.c_int64
    lib.pam_context_get_tuning.argtypes = [C.c_void_p, C.c_char_p]
    lib.pam_calc_tmatrix.restype = C.c_int
    lib.pam_calc_tmatrix.argtypes = ([C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32,
                                      C.c_int32] + [_dp] * 9 + [_ip])
    return lib


EXPORTED_SYMBOLS = ["pam_version", "pam_device_count", "pam_context_create",
                    "pam_context_destroy", "pam_run", "pam_run_device", "pam_run_device_noalloc",
                    "pam_last_stage_times", "pam_stage_names", "pam_last_work",
                    "pam_workspace_bytes", "pam_context_set_workspace_limit", "pam_context_bytes",
                    "pam_context_trim", "pam_tuning_keys", "pam_context_set_tuning",
                    "pam_context_get_tuning", "pam_nummu", "pam_calc_tmatrix"]

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(PACKAGE_DIR, LIBRARIES[NUMMU])


def library_path(nummu: int = NUMMU) -> str:
    if nummu not in LIBRARIES:
        raise NotImplementedError(
            f"{nummu} quadrature angles: the library is built for {sorted(LIBRARIES)}")
    return os.path.join(PACKAGE_DIR, LIBRARIES[nummu])


def nangles(sets) -> int:
    """Angles per hemisphere of a run: pyPamtra's _nangles (core.py:263),
    passed as sets["_nangles"]; 16 by default."""
    n = int((sets or {}).get("_nangles", NUMMU))
    library_path(n)
    return n
