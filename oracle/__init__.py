"""Loader of the CPU oracle (oracle/pam_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by the cpu_baseline leg of bench.py.  The product package
``pamtra_amd`` never imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from pamtra_amd import _capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpam_oracle.so")
LIB_PATHS = {16: LIB_PATH, 32: os.path.join(HERE, "libpam_oracle_n64.so")}
_libs = {}
MAXNLEG = 200  # PAM_MAXNLEG of include/pamtra_amd.h
# the 32-angle oracle keeps 64x64 operators on the stack of its OpenMP
# threads; libgomp reads this when it starts
os.environ.setdefault("OMP_STACKSIZE", "64M")


def build(force: bool = False) -> str:
    """Compiles the oracle (both angle counts) with the committed Makefile."""
    if force or not all(os.path.exists(p) for p in LIB_PATHS.values()):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    from . import refbind
    refbind.build(force=force)
    return LIB_PATH


def lib(nummu: int = 16) -> C.CDLL:
    if nummu not in _libs:
        path = LIB_PATHS[nummu]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.pam_oracle_run.restype = C.c_int
        L.pam_oracle_run.argtypes = [C.POINTER(_capi.PamSettings), C.POINTER(_capi.PamHydroDesc),
                                     C.POINTER(_capi.PamHydro4D), C.POINTER(_capi.PamProfile),
                                     dp, C.c_int, C.POINTER(_capi.PamResults), C.c_int,
                                     C.POINTER(C.c_int64)]
        L.pam_oracle_lobatto.argtypes = [C.c_int, dp, dp]
        L.pam_oracle_gausquad.argtypes = [C.c_int, dp, dp]
        L.pam_oracle_eps_water_tkc.argtypes = [C.c_double, C.c_double, dp, dp]
        L.pam_oracle_eps_ice.argtypes = [C.c_double, C.c_double, dp, dp]
        L.pam_oracle_gasabs.restype = C.c_double
        L.pam_oracle_gasabs.argtypes = [C.c_double] * 4
        L.pam_oracle_gasabs_parts.argtypes = [C.c_double] * 4 + [dp, dp]
        L.pam_oracle_fastem_parts.argtypes = [C.c_double] * 5 + [dp]
        L.pam_oracle_eps_mix.argtypes = [C.c_double] * 3 + [dp, dp]
        fp = C.POINTER(C.c_float)
        L.pam_oracle_tm_tmatrix.restype = C.c_int
        L.pam_oracle_tm_tmatrix.argtypes = [C.c_double] * 5 + [fp, C.c_int]
        L.pam_oracle_tm_single.restype = C.c_int
        L.pam_oracle_tm_single.argtypes = [C.c_double] * 6 + [dp] * 4
        L.pam_oracle_calc_tmatrix.restype = C.c_int
        L.pam_oracle_calc_tmatrix.argtypes = [C.c_double] * 3 + [C.c_int, C.c_int] + [dp] * 10
        L.pam_oracle_mie_coefs.restype = C.c_int
        L.pam_oracle_mie_coefs.argtypes = [C.c_double] * 3 + [C.c_int, C.POINTER(C.c_int), dp, dp,
                                                              dp]
        L.pam_oracle_mie_spheres.restype = C.c_int
        L.pam_oracle_mie_spheres.argtypes = [C.c_double, C.c_int, C.c_int, dp, dp, dp, dp,
                                             C.c_double, C.c_double, dp, dp, C.POINTER(C.c_int)]
        L.pam_oracle_tessem2.argtypes = [C.c_double] * 5 + [dp, dp]
        L.pam_oracle_fastem.argtypes = [C.c_double] * 5 + [dp, dp]
        L.pam_oracle_liu_permittivity.argtypes = [C.c_double] * 3 + [dp, dp]
        L.pam_oracle_mie_single.restype = C.c_int
        L.pam_oracle_mie_single.argtypes = [C.c_double, C.c_double, C.c_double,
                                            C.POINTER(C.c_int), dp, dp, dp]
        L.pam_oracle_mie_angle.argtypes = [C.c_double] * 4 + [dp]
        L.pam_oracle_layer_doubling.restype = C.c_int
        L.pam_oracle_layer_doubling.argtypes = [dp] + [C.c_double] * 7 + [dp, dp, dp, dp,
                                                                          C.POINTER(C.c_int)]
        L.pam_oracle_ssrg_shape.restype = C.c_double
        L.pam_oracle_ssrg_shape.argtypes = [C.c_double] * 5
        L.pam_oracle_ssrg.restype = C.c_int
        L.pam_oracle_ssrg.argtypes = [C.c_double, C.c_int, dp, dp, dp, dp] + [C.c_double] * 8 + \
            [dp, dp, C.POINTER(C.c_int)]
        L.pam_oracle_scatcnv.argtypes = [C.c_int, dp, C.c_double, C.c_double, dp]
        L.pam_oracle_esat.restype = C.c_double
        L.pam_oracle_esat.argtypes = [C.c_double]
        L.pam_oracle_planck.restype = C.c_double
        L.pam_oracle_planck.argtypes = [C.c_double, C.c_double]
        L.pam_oracle_rad2tb.restype = C.c_double
        L.pam_oracle_rad2tb.argtypes = [C.c_double, C.c_double]
        L.pam_oracle_dsd.restype = C.c_int
        L.pam_oracle_dsd.argtypes = [C.POINTER(_capi.PamHydroDesc), C.POINTER(_capi.PamSettings),
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, dp, dp, dp, dp, dp]
        _libs[nummu] = L
    return _libs[nummu]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def run(nmlSets, sets, descriptorFile, descriptorFile4D, profile, freqs, nthreads=1,
        diagnostics=False, descriptorFileFS=None):
    """CPU forward operator on the same inputs as the product's
    PamtraFortranWrapper.  Returns (results dict, status, work counters)."""
    m = _capi.marshal(nmlSets, sets, descriptorFile, descriptorFile4D, profile, descriptorFileFS)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f), diagnostics=diagnostics)
    work = (C.c_int64 * 4)()
    st = lib(m.nummu).pam_oracle_run(C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                              C.byref(m.profile), fp, len(f), C.byref(hr.struct),
                              int(nthreads), work)
    out = dict(hr.arrays)
    out["msg"] = _capi.results_message(hr)
    return out, st, list(work)


def lobatto(num=16):
    mu, w = np.zeros(num), np.zeros(num)
    lib(num if num in LIB_PATHS else 16).pam_oracle_lobatto(num, _dp(mu), _dp(w))
    return mu, w


def gausquad(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().pam_oracle_gausquad(n, _dp(x), _dp(w))
    return x, w


def eps_water_tkc(t_c, f_ghz):
    re, im = C.c_double(), C.c_double()
    lib().pam_oracle_eps_water_tkc(t_c, f_ghz, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def eps_ice(t_k, f_ghz):
    re, im = C.c_double(), C.c_double()
    lib().pam_oracle_eps_ice(t_k, f_ghz, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def gasabs(f_ghz, t_k, rho_v, p_pa):
    return lib().pam_oracle_gasabs(f_ghz, t_k, rho_v, p_pa)


def gasabs_parts(f_ghz, t_k, rho_v, p_pa):
    """(dry air, water vapour) absorption in Np/km."""
    a, w = C.c_double(), C.c_double()
    lib().pam_oracle_gasabs_parts(f_ghz, t_k, rho_v, p_pa, C.byref(a), C.byref(w))
    return a.value, w.value


def fastem_parts(f, zen, temp, sal, wind):
    """(Re eps, Im eps, Rv, Rh Fresnel, foam cover, Rv, Rh foam, Rv, Rh
    large-scale, small-scale factor, cos zenith)."""
    out = np.zeros(11)
    lib().pam_oracle_fastem_parts(f, zen, temp, sal, wind, _dp(out))
    return out


def eps_mix(eps1, mix_de):
    re, im = C.c_double(), C.c_double()
    lib().pam_oracle_eps_mix(eps1.real, eps1.imag, mix_de, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def mie_coefs(x, m, maxn=10000):
    """(status, nterms, a[n], b[n], (Qext, Qsca, Qback)); m = n + i k, k >= 0."""
    a, b, q = np.zeros((maxn, 2)), np.zeros((maxn, 2)), np.zeros(3)
    nt = C.c_int()
    st = lib().pam_oracle_mie_coefs(x, m.real, abs(m.imag), maxn, C.byref(nt), _dp(a), _dp(b),
                                    _dp(q))
    n = min(nt.value, maxn)
    return st, nt.value, a[:n, 0] + 1j * a[:n, 1], b[:n, 0] + 1j * b[:n, 1], q


def mie_spheres(freq_hz, liq_ice, diameter, del_d, ndens, density, m):
    """(status, (ext, albedo, back), legen[3][200], nlegen)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (diameter, del_d, ndens, density)]
    out, legen = np.zeros(3), np.zeros((3, MAXNLEG))
    nl = C.c_int()
    st = lib().pam_oracle_mie_spheres(freq_hz, liq_ice, len(arrs[0]), *[_dp(a) for a in arrs],
                                      m.real, abs(m.imag), _dp(out), _dp(legen), C.byref(nl))
    return st, out, legen, nl.value


def tessem2(f, zen, wind, temp, sal):
    e, r = np.zeros(2), np.zeros(2)
    lib().pam_oracle_tessem2(f, zen, wind, temp, sal, _dp(e), _dp(r))
    return e, r


def fastem(f, zen, temp, sal, wind):
    e, r = np.zeros(2), np.zeros(2)
    lib().pam_oracle_fastem(f, zen, temp, sal, wind, _dp(e), _dp(r))
    return e, r


def liu_permittivity(temp, sal, f):
    re, im = C.c_double(), C.c_double()
    lib().pam_oracle_liu_permittivity(temp, sal, f, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def mie_single(x, m):
    nt = C.c_int()
    qe, qs, qb = C.c_double(), C.c_double(), C.c_double()
    st = lib().pam_oracle_mie_single(x, m.real, -m.imag if m.imag < 0 else m.imag,
                                     C.byref(nt), C.byref(qe), C.byref(qs), C.byref(qb))
    assert st == 0
    return nt.value, qe.value, qs.value, qb.value


def mie_angle(x, m, mu):
    p = np.zeros(4)
    lib().pam_oracle_mie_angle(x, m.real, abs(m.imag), mu, _dp(p))
    return p


def scatcnv(nlegen, coef, kext, salb, nummu=16):
    coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(3, nlegen + 1)
    n = 2 * nummu
    smat = np.zeros((4, n, n))
    lib(nummu).pam_oracle_scatcnv(nlegen, _dp(coef), kext, salb, _dp(smat))
    return smat


def layer_doubling(smat, kext, salb, gas, zdiff, t_top, t_bot, wavelength_um):
    smat = np.ascontiguousarray(smat, dtype=np.float64)
    n = smat.shape[-1]
    R, T = np.zeros((n, n)), np.zeros((n, n))
    sp, sm = np.zeros(n), np.zeros(n)
    nd = C.c_int()
    st = lib(n // 2).pam_oracle_layer_doubling(_dp(smat), kext, salb, gas, zdiff, t_top, t_bot,
                                         wavelength_um, _dp(R), _dp(T), _dp(sp), _dp(sm),
                                         C.byref(nd))
    return st, R, T, sp, sm, nd.value


def ssrg_shape(x, kappa, beta, gamma, zeta):
    """Self-similar Rayleigh-Gans shape factor at electrical size x."""
    return lib().pam_oracle_ssrg_shape(x, kappa, beta, gamma, zeta)


def ssrg(freq_hz, diameter, del_d, ndens, mass, as_ratio, canting, m, kappa, beta, gamma, zeta):
    """One SSRGA hydrometeor (Legendre-series form).  m: refractive index of
    solid ice, n + i k with k >= 0.  Returns (status, extinction, albedo,
    back_scatt, legen[3][MAXNLEG], nlegen)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (diameter, del_d, ndens, mass)]
    out = np.zeros(3)
    legen = np.zeros((3, MAXNLEG))
    nl = C.c_int()
    st = lib().pam_oracle_ssrg(freq_hz, len(arrs[0]), *[_dp(a) for a in arrs], as_ratio, canting,
                               m.real, m.imag, kappa, beta, gamma, zeta, _dp(out), _dp(legen),
                               C.byref(nl))
    return st, out[0], out[1], out[2], legen, nl.value


def tm_tmatrix(lam, m, axi, as_ratio, cap=48):
    """The stored (REAL*4) T matrix of one spheroid: (nmax, t[8][cap+1][cap][cap])
    with components RT11, RT12, RT21, RT22, IT11, IT12, IT21, IT22 indexed
    [m][n-1][n'-1].  m = n + i k (EBCM convention), lam and axi in metres."""
    t = np.zeros((8, cap + 1, cap, cap), np.float32)
    nm = lib().pam_oracle_tm_tmatrix(lam, m.real, abs(m.imag), axi, as_ratio,
                                     t.ctypes.data_as(C.POINTER(C.c_float)), cap)
    return nm, t


def calc_tmatrix(freq_hz, m, phase, dmax, del_d, ndens, density, as_ratio, canting, nummu=16):
    """calc_tmatrix over the bins of one hydrometeor.  m = n - i k, the
    refractive index scatProperties passes.  Returns (status, smat, ext, emis,
    sback) in the reference's Fortran shapes (2,nummu,2,nummu,2), (2,2,nummu),
    (2,nummu), (2,2)."""
    arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, float), (len(dmax),)))
            for a in (dmax, del_d, ndens, density, as_ratio, canting)]
    sm, ex = np.zeros(2 * nummu * 2 * nummu * 2), np.zeros(4 * nummu)
    em, sb = np.zeros(2 * nummu), np.zeros(4)
    st = lib(nummu).pam_oracle_calc_tmatrix(freq_hz, m.real, -m.imag, phase, len(dmax),
                                            *[_dp(a) for a in arrs], _dp(sm), _dp(ex), _dp(em),
                                            _dp(sb))
    return (st, sm.reshape((2, nummu, 2, nummu, 2), order="F"),
            ex.reshape((2, 2, nummu), order="F"), em.reshape((2, nummu), order="F"),
            sb.reshape((2, 2), order="F"))


def esat(t_k):
    """Goff-Gratch saturation vapour pressure over water [hPa]."""
    return lib().pam_oracle_esat(t_k)


def planck(t, wl):
    return lib().pam_oracle_planck(t, wl)


def rad2tb(r, wl):
    return lib().pam_oracle_rad2tb(r, wl)


def dsd(desc_row: dict, nmlSets: dict, q_h, n_tot, r_eff, layer_t, press=8.5e4, sfc_type=0):
    """run_drop_size_dist of one (layer, hydrometeor): (status, [d_ds,
    delta_d_ds, n_ds, density2scat, diameter2scat]).  press and sfc_type
    enter the ECHAM monodisperse forms only."""
    rec = np.rec.fromrecords([tuple(desc_row[k] for k in DESC_NAMES)], names=DESC_NAMES)
    h = _capi.make_hydros(rec)
    s = _capi.make_settings(nmlSets)
    nb = int(desc_row["nbin"]) if desc_row["dist_name"] not in ("mono", "mono_cosmo_ice") else 2
    arrs = [np.zeros(nb) for _ in range(5)]
    st = lib().pam_oracle_dsd(h, C.byref(s), q_h, n_tot, r_eff, layer_t, press, int(sfc_type),
                              *[_dp(a) for a in arrs])
    return st, arrs


DESC_NAMES = ["hydro_name", "as_ratio", "liq_ice", "rho_ms", "a_ms", "b_ms", "alpha_as",
              "beta_as", "moment_in", "nbin", "dist_name", "p_1", "p_2", "p_3", "p_4", "d_1",
              "d_2", "scat_name", "vel_size_mod", "canting"]
