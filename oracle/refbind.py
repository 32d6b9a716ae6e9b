"""Loader of oracle/_ref/libpam_ref.so: routines of the reference's own
Fortran sources compiled unchanged (oracle/ref.mk) behind the C-callable
driver oracle/pam_ref_driver.f90.

TEST INFRASTRUCTURE ONLY: the tests use it to pin the C oracle against the
reference's compiled code.  It exists only where the reference sources were
present at build time; callers skip when ``available()`` is False.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libpam_ref.so")
REF_SRC = "/root/reference/src"
MAXNLEG = 200
_lib = None


def build(force: bool = False) -> str | None:
    """Builds the library when the reference sources are present."""
    if not os.path.isdir(REF_SRC):
        return LIB_PATH if os.path.exists(LIB_PATH) else None
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-f", os.path.join(HERE, "ref.mk")] + (["-B"] if force else []),
                       check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def available() -> bool:
    try:
        return build() is not None and os.path.exists(LIB_PATH)
    except (OSError, subprocess.CalledProcessError):
        return False


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        # lazy binding: one routine of vars_atmosphere.f90 that is never called
        # references GMTIME, which flang's runtime does not provide
        L = C.CDLL(LIB_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.ref_gasabs.argtypes = [C.c_double] * 4 + [dp, dp, ip]
        L.ref_tessem2.argtypes = [C.c_double] * 5 + [dp, dp]
        L.ref_fastem_parts.argtypes = [C.c_double] * 5 + [dp]
        L.ref_eps_water_tkc.argtypes = [C.c_double] * 2 + [dp, dp]
        L.ref_eps_ice.argtypes = [C.c_double] * 2 + [dp, dp]
        L.ref_eps_mix.argtypes = [C.c_double] * 3 + [dp, dp]
        L.ref_esat.restype = C.c_double
        L.ref_esat.argtypes = [C.c_double]
        L.ref_q2abs.restype = C.c_double
        L.ref_q2abs.argtypes = [C.c_double] * 5
        L.ref_lobatto.argtypes = [C.c_int, dp, dp]
        L.ref_gausquad.argtypes = [C.c_int, dp, dp]
        L.ref_mie.restype = C.c_int
        L.ref_mie.argtypes = [C.c_double] * 3 + [C.c_int, ip, dp, dp, dp]
        L.ref_mie_angle.argtypes = [C.c_double] * 4 + [dp]
        L.ref_mie_spheres.restype = C.c_int
        L.ref_mie_spheres.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, dp, dp,
                                      C.c_double, C.c_double, dp, dp, dp, ip]
        L.ref_scatcnv.argtypes = [C.c_int, dp, C.c_double, C.c_double, dp, dp, dp]
        L.ref_calc_tmatrix.restype = C.c_int
        L.ref_calc_tmatrix.argtypes = ([C.c_double] * 3 + [C.c_int, C.c_int] + [dp] * 6 +
                                       [C.c_double] + [dp] * 4)
        L.ref_tmatrix_amplq.restype = C.c_int
        L.ref_tmatrix_amplq.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_float), C.c_int]
        L.ref_dsd.restype = C.c_int
        L.ref_dsd.argtypes = [C.c_char_p, dp, ip, dp, ip] + [dp] * 7
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def gasabs(f_ghz, t_k, rho_v, p_pa):
    """(dry air, water vapour) absorption in Np/km and the error status."""
    a, w, e = C.c_double(), C.c_double(), C.c_int()
    lib().ref_gasabs(f_ghz, t_k, rho_v, p_pa, C.byref(a), C.byref(w), C.byref(e))
    return a.value, w.value, e.value


def tessem2(f, zen, wind, temp, sal):
    e, r = np.zeros(2), np.zeros(2)
    lib().ref_tessem2(f, zen, wind, temp, sal, _dp(e), _dp(r))
    return e, r


def fastem_parts(f, zen, temp, sal, wind):
    out = np.zeros(11)
    lib().ref_fastem_parts(f, zen, temp, sal, wind, _dp(out))
    return out


def _cplx(fn, *args):
    re, im = C.c_double(), C.c_double()
    fn(*args, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def eps_water_tkc(t_c, f_ghz):
    return _cplx(lib().ref_eps_water_tkc, t_c, f_ghz)


def eps_ice(t_k, f_ghz):
    return _cplx(lib().ref_eps_ice, t_k, f_ghz)


def eps_mix(eps1, mix_de):
    return _cplx(lib().ref_eps_mix, eps1.real, eps1.imag, mix_de)


def esat(t_k):
    return lib().ref_esat(t_k)


def q2abs(spec, t, p, qv, qall):
    return lib().ref_q2abs(spec, t, p, qv, qall)


def lobatto(n):
    mu, w = np.zeros(n), np.zeros(n)
    lib().ref_lobatto(n, _dp(mu), _dp(w))
    return mu, w


def gausquad(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().ref_gausquad(n, _dp(x), _dp(w))
    return x, w


def mie(x, m, maxn=10000):
    """miecalc + miecross: (status, nterms, a[n], b[n], (Qext, Qsca, Qback)).
    m = n + i k with k >= 0."""
    a, b, q = np.zeros((maxn, 2)), np.zeros((maxn, 2)), np.zeros(3)
    nt = C.c_int()
    st = lib().ref_mie(x, m.real, abs(m.imag), maxn, C.byref(nt), _dp(a), _dp(b), _dp(q))
    n = min(nt.value, maxn)
    return st, nt.value, a[:n, 0] + 1j * a[:n, 1], b[:n, 0] + 1j * b[:n, 1], q


def mie_angle(x, m, mu):
    p = np.zeros(4)
    lib().ref_mie_angle(x, m.real, abs(m.imag), mu, _dp(p))
    return p


def mie_spheres(freq_hz, t, liq_ice, diameter, del_d, ndens, density, m):
    """calc_mie_spheres: (status, (ext, albedo, back), legen[4][200], back_spec, nlegen)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (diameter, del_d, ndens, density)]
    n = len(arrs[0])
    out, legen, bs = np.zeros(3), np.zeros((4, MAXNLEG)), np.zeros(n)
    nl = C.c_int()
    st = lib().ref_mie_spheres(freq_hz, t, liq_ice, n, *[_dp(a) for a in arrs], m.real,
                               abs(m.imag), _dp(out), _dp(legen), _dp(bs), C.byref(nl))
    return st, out, legen, bs, nl.value


def tmatrix_amplq(lam, m, axi, as_ratio, cap=48):
    """tmatrix_amplq for one spheroid: (nmax, t) as oracle.tm_tmatrix."""
    t = np.zeros((8, cap + 1, cap, cap), np.float32)
    nm = lib().ref_tmatrix_amplq(lam, m.real, abs(m.imag), axi, as_ratio,
                                 t.ctypes.data_as(C.POINTER(C.c_float)), cap)
    return nm, t


def calc_tmatrix(freq_hz, m, phase, dmax, del_d, ndens, density, as_ratio, canting, temp=273.15):
    """calc_tmatrix (tmatrix_db 'none', passive) as oracle.calc_tmatrix."""
    arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, float), (len(dmax),)))
            for a in (dmax, del_d, ndens, density, as_ratio, canting)]
    sm, ex, em, sb = np.zeros(2 * 16 * 2 * 16 * 2), np.zeros(64), np.zeros(32), np.zeros(4)
    st = lib().ref_calc_tmatrix(freq_hz, m.real, -m.imag, phase, len(dmax),
                                *[_dp(a) for a in arrs], temp, _dp(sm), _dp(ex), _dp(em),
                                _dp(sb))
    return (st, sm.reshape((2, 16, 2, 16, 2), order="F"), ex.reshape((2, 2, 16), order="F"),
            em.reshape((2, 16), order="F"), sb.reshape((2, 2), order="F"))


def scatcnv(nlegen, coef3, kext, salb, nummu=16):
    """scatcnv with the coefficient rows of calc_scatProperties (P11, P12,
    P33, P34 = 0, P11, P33).  Returns the phase matrix in the oracle's layout
    [quadrant][(j2, i2)][(j1, i1)] and the Fortran ext_matrix, emis_vec."""
    coef = np.zeros((MAXNLEG, 6))  # Fortran coef(6, 200): index [l][row] in C order
    c3 = np.asarray(coef3, dtype=np.float64).reshape(3, nlegen + 1)
    coef[:nlegen + 1, 0] = c3[0]
    coef[:nlegen + 1, 1] = c3[1]
    coef[:nlegen + 1, 2] = c3[2]
    coef[:nlegen + 1, 4] = c3[0]
    coef[:nlegen + 1, 5] = c3[2]
    smat = np.zeros(2 * nummu * 2 * nummu * 4)
    ext = np.zeros(2 * 2 * nummu * 2)
    emis = np.zeros(2 * nummu * 2)
    lib().ref_scatcnv(nlegen, _dp(coef), kext, salb, _dp(smat), _dp(ext), _dp(emis))
    # Fortran scatter_matrix(i2, j2, i1, j1, q) -> C order [q][j1][i1][j2][i2]
    s = smat.reshape(4, nummu, 2, nummu, 2)
    n = 2 * nummu
    out = np.empty((4, n, n))
    for q in range(4):
        # rows (j2, i2) = i2 + 2 j2, columns (j1, i1) = i1 + 2 j1
        out[q] = s[q].transpose(2, 3, 0, 1).reshape(n, n)
    return out, ext, emis


DSD_REAL = ["as_ratio", "rho_ms", "a_ms", "b_ms", "alpha_as", "beta_as", "p_1", "p_2", "p_3",
            "p_4", "d_1", "d_2", "canting"]


def dsd(row: dict, q_h, n_tot, r_eff, layer_t, press=8.5e4, sfc_type=0, adaptive=False,
        conserve=True, limit_density_area=True, softsphere_min_density=10.0):
    """run_drop_size_dist for one (layer, hydrometeor).  Returns (status,
    dict of d_ds, delta_d_ds, n_ds, density2scat, diameter2scat, mass_ds,
    area_ds)."""
    dreal = np.array([float(row[k]) for k in DSD_REAL] + [softsphere_min_density])
    dint = (C.c_int * 7)(int(row["liq_ice"]), int(row["moment_in"]), int(row["nbin"]),
                         int(adaptive), int(conserve), int(limit_density_area), int(sfc_type))
    layer = np.array([q_h, n_tot, r_eff, layer_t, press], dtype=np.float64)
    nmax = max(int(row["nbin"]), 2) + 1
    outs = [np.zeros(nmax) for _ in range(7)]
    nb = C.c_int()
    st = lib().ref_dsd(str(row["dist_name"]).encode()[:15], _dp(dreal), dint, _dp(layer),
                       C.byref(nb), *[_dp(o) for o in outs])
    n = nb.value
    keys = ["d_ds", "delta_d_ds", "n_ds", "density2scat", "diameter2scat", "mass_ds", "area_ds"]
    return st, {k: o[:n] for k, o in zip(keys, outs)}
