"""Every size distribution the device builds, and every moment_in branch of
make_dist_params (make_dist_param.f90:2-631), through the HIP path against
the oracle (whose distribution chain is pinned against the reference's own
Fortran in test_ref_pins.py).

One run holds one hydrometeor per case, each present in its own column
(moment_in 0 cases are present everywhere, as in hydrometeor_extinction.f90).
Compared: TB within 1e-3 K, per-layer extinction kexttot within 1e-9
relative, and the Legendre orders exactly; fixed and adaptive grids."""

import ctypes as C

import numpy as np
import pytest

import oracle
import pamtra_amd
from pamtra_amd import _capi

FREQS = [23.8, 89.0]


def row(name, liq=1, mi=3, nbin=40, dist="exp", p=(-99.0, -99.0, -99.0, -99.0), d=(1e-6, 6e-3),
        a_ms=-99.0, b_ms=-99.0, rho_ms=-99.0, as_ratio=-99.0):
    return (name, as_ratio, liq, rho_ms, a_ms, b_ms, -99.0, -99.0, mi, nbin, dist, *p, *d,
            "mie-sphere", "khvorostyanov01_drops", -99.0)


# (descriptor row, q [kg/kg], n [#/kg], r_eff [m]) per case
CASES = [
    (row("mono_d", dist="mono", d=(2e-5, -99.0)), 2e-4, None, None),
    (row("mono_n", dist="mono", mi=1, d=(3e-5, -99.0)), None, 1e8, None),
    (row("mono_p1", dist="mono", p=(1e8, -99.0, -99.0, -99.0), d=(-99.0, -99.0)), 3e-4, None,
     None),
    (row("mono_23", dist="mono", mi=23, d=(-99.0, -99.0)), 2e-4, None, 1.5e-5),
    (row("cosmo_ice", liq=-1, dist="mono_cosmo_ice", a_ms=130.0, b_ms=3.0, d=(-99.0, -99.0)), 5e-5,
     None, None),
    (row("echam_cl", dist="mono_echam_cl", nbin=2, d=(-99.0, -99.0)), 3e-4, None, None),
    (row("echam_ice", liq=-1, dist="mono_echam_ice", nbin=2, rho_ms=500.0, d=(-99.0, -99.0)),
     4e-5, None, None),
    (row("exp_n0", p=(-99.0, -99.0, 8e6, -99.0)), 5e-4, None, None),
    (row("exp_lam_n", mi=1, p=(-99.0, 3e-4, -99.0, -99.0)), None, 2e3, None),
    (row("exp_13", mi=13), 5e-4, 3e3, None),
    (row("exp_23", mi=23), 5e-4, None, 4e-4),
    (row("exp_12", mi=12), None, 3e3, 4e-4),
    (row("exp_ryan", liq=-1, dist="exp_ryan", rho_ms=400.0, d=(1e-6, 1e-2)), 1e-4, None, None),
    (row("cosmo_snow", liq=-1, dist="exp_cosmo_snow", a_ms=0.038, b_ms=2.0, d=(0.51e-10, 2e-2)),
     2e-4, None, None),
    (row("mgam_13", dist="mgamma", mi=13, p=(-99.0, -99.0, 2.0, 1.0), d=(1e-6, 1e-4)), 3e-4, 1e8,
     None),
    (row("mgam_12", dist="mgamma", mi=12, p=(-99.0, -99.0, 2.0, 1.0), d=(1e-6, 1e-4)), None, 1e8,
     1e-5),
    (row("mgam_p2", dist="mgamma", p=(-99.0, 1e-5, 2.0, 1.0), d=(1e-6, 1e-4)), 3e-4, None, None),
    (row("mnh_rain", dist="mgamma_MNH", p=(1e7, -1.0, 0.0, 1.0), d=(1.2e-4, 6e-3)), 4e-4, None,
     None),
    (row("mnh_snow", liq=-1, dist="mgamma_MNH", p=(5.0, 1.0, 0.0, 1.0), a_ms=0.02, b_ms=1.9,
         d=(0.51e-10, 2e-2)), 2e-4, None, None),
    (row("logn_3", dist="logn", p=(1e8, -99.0, 0.3, -99.0), d=(1e-6, 1e-4)), 3e-4, None, None),
    (row("logn_23", dist="logn", mi=23, p=(-99.0, -99.0, 0.3, -99.0), d=(1e-6, 1e-4)), 3e-4, None,
     1e-5),
    (row("ngam_23", dist="norm_gamma", mi=23, p=(-99.0, -99.0, 1.5, -99.0)), 5e-4, None, 5e-4),
    (row("ngam_rain", dist="norm_gamma", mi=23), 5e-4, None, 5e-4),
    (row("ngam_0", dist="norm_gamma", mi=0, p=(1e-3, 8e6, 2.0, -99.0)), None, None, None),
    (row("nmgam_0", dist="norm_mgamma", mi=0, p=(1e-3, 8e6, 1.0, 1.5)), None, None, None),
]


def build(cases, land=False):
    pam = pamtra_amd.pyPamtra()
    for c in cases:
        pam.df.addHydrometeor(c[0])
    ncol, L = len(cases), 3
    hl = np.array([0.0, 1000.0, 2000.0, 3000.0])
    _, p, t = pamtra_amd.meteoSI.usStandard(hl, geopotential=False)
    tile = lambda a: np.tile(a, (ncol, 1))  # noqa: E731
    pam.createProfile(hgt_lev=tile(hl), press_lev=tile(p), temp_lev=tile(t),
                      relhum_lev=tile(np.array([70.0, 60.0, 50.0, 40.0])))
    pam.set["verbose"] = 0
    pam.set["pyVerbose"] = 0
    nh = len(cases)
    q = np.zeros((ncol, 1, L, nh))
    n = np.full((ncol, 1, L, nh), -9999.0)
    r = np.full((ncol, 1, L, nh), -9999.0)
    for k, (_, qk, nk, rk) in enumerate(cases):
        for z in (0, 1):  # two layers of each column, scaled
            if qk is not None:
                q[k, 0, z, k] = qk * (1 + z)
            if nk is not None:
                n[k, 0, z, k] = nk * (1 + 0.5 * z)
            if rk is not None:
                r[k, 0, z, k] = rk
    pam.p["hydro_q"], pam.p["hydro_n"], pam.p["hydro_reff"] = q, n, r
    if land:
        # any sfc_type above 0.5 is land to the ECHAM forms
        # (make_dist_param.f90:183); 2 keeps the constant-emissivity surface
        # of sfc_optics.f90:79, where 1 would need a land-emissivity atlas
        pam.p["sfc_type"][:] = 2
    return pam


def run_both(pam, adaptive):
    from pamtra_amd import _lib
    nml = dict(pam.nmlSet)
    nml["active"] = False
    nml["hydro_adaptive_grid"] = adaptive
    lib, ctx = _lib.library(), _lib.get_context(0)
    m = _capi.marshal(nml, pam.set, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(FREQS)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    print(f"pam_run dsd cases={len(pam.df.data)} adaptive={adaptive}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    g = dict(hr.arrays)
    assert st == 0, _capi.results_message(hr)
    c, st2, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, FREQS, nthreads=8,
                           diagnostics=True)
    assert st2 == 0
    return g, c


@pytest.mark.gpu
@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("land", [False, True])
def test_every_distribution_vs_oracle(adaptive, land):
    g, c = run_both(build(CASES, land), adaptive)
    d = np.abs(g["tb"] - c["tb"])
    print(f"max |dTB| = {d.max():.3e} K", flush=True)
    assert np.all(np.isfinite(g["tb"]))
    assert d.max() <= 1e-3
    np.testing.assert_array_equal(g["nlegen"], c["nlegen"])
    np.testing.assert_allclose(g["kexttot"], c["kexttot"], rtol=1e-9, atol=1e-300)
    # every case scatters: its column holds extinction beyond the gas
    assert np.all(np.nanmax(c["kexttot"][:, 0, :, 0], axis=1) > 0)


def test_oracle_runs_every_case():
    """CPU check of the scenario itself: every case is valid and scatters."""
    pam = build(CASES)
    nml = dict(pam.nmlSet)
    nml["active"] = False
    c, st, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, FREQS[:1], nthreads=8,
                          diagnostics=True)
    assert st == 0
    assert np.all(np.nanmax(c["kexttot"][:, 0, :, 0], axis=1) > 0)


def test_refused_distributions():
    """'const' and 'const_cosmo_ice' leave the reference's bins unset
    (make_dist.f90:164-215); ECHAM monodisperse rows need nbin = 2."""
    for dist in ("const", "const_cosmo_ice"):
        pam = build([(row("c", dist=dist), 1e-4, None, None)])
        with pytest.raises(NotImplementedError, match="leaves its bins unset"):
            _capi.make_hydros(pam.df.data)
    pam = build([(row("e", dist="mono_echam_cl", nbin=5), 1e-4, None, None)])
    with pytest.raises(ValueError, match="nbin = 2"):
        _capi.make_hydros(pam.df.data)
