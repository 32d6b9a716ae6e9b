"""Self-similar Rayleigh-Gans aggregates in Legendre-series form, the
"ssrg-rt3" scattering name (reference scatProperties.f90:771-863,
rayleigh_gans.f90:467-872).

CPU: the oracle pinned by the Fourier structure the shape factor stands for
(Hogan and Westbrook 2014; Hogan et al. 2017 eq. 4), the Rayleigh limit of
the cross sections and of the Legendre series, the reconstruction of the
phase function from its series, and the parsing of the scattering name.
GPU: the device producer against the oracle through the whole forward
operator (FP64 bar |dTB| <= 1e-3 K; mixed <= 0.05 K), reflectivities, and
the inputs the reference refuses."""

import ctypes as C

import numpy as np
import pytest
from scipy.integrate import quad

import oracle
import pamtra_amd
from pamtra_amd import _capi, importer, workloads
from _scenario import passive_nml

C_LIGHT = 299792458.0


# ---------------------------------------------------------------- oracle pins


@pytest.mark.parametrize("kappa", [0.0, 0.19, 0.25])
@pytest.mark.parametrize("x", [0.3, 1.0, 2.5, 5.0, 12.0])
def test_shape_factor_mean_structure_is_its_fourier_power(kappa, x):
    """With beta = 0 the shape factor is (pi^2/4) |F(x)|^2, F the Fourier
    integral of the mean structure A(s) = (1 + kappa/3) cos(pi s) +
    kappa cos(3 pi s) over the particle, s in [-1/2, 1/2] (F(0) = 2/pi)."""
    A = lambda s: (1.0 + kappa / 3.0) * np.cos(np.pi * s) + kappa * np.cos(3.0 * np.pi * s)
    F = quad(lambda s: A(s) * np.cos(2.0 * x * s), -0.5, 0.5, epsabs=1e-14, epsrel=1e-13)[0]
    got = oracle.ssrg_shape(x, kappa, 0.0, 1.66, 1.0)
    np.testing.assert_allclose(got, np.pi ** 2 / 4.0 * F * F, rtol=1e-10)


def _harmonic_power(j, x):
    """Phase-averaged |Fourier integral|^2 of cos(2 pi j s + phi) over the
    particle: the fluctuation spectrum's j-th term, by quadrature."""
    def power(phi):
        re = quad(lambda s: np.cos(2 * np.pi * j * s + phi) * np.cos(2 * x * s), -0.5, 0.5)[0]
        im = quad(lambda s: np.cos(2 * np.pi * j * s + phi) * np.sin(2 * x * s), -0.5, 0.5)[0]
        return re * re + im * im
    return quad(power, 0.0, 2.0 * np.pi)[0] / (2.0 * np.pi)


@pytest.mark.parametrize("x", [0.7, 4.1, 9.3])
def test_shape_factor_fluctuations_are_harmonic_powers(x):
    """The beta term is (pi^2/4) beta sum_j (2j)^-gamma zeta_j P_j(x), P_j the
    phase-averaged power of the j-th harmonic, zeta_1 = zeta, zeta_j = 1,
    j up to floor(5x/pi + 1)."""
    kappa, beta, gamma, zeta = 0.19, 0.23, 5.0 / 3.0, 0.4
    jmax = int(np.floor(5.0 * x / np.pi + 1.0))
    want = sum((zeta if j == 1 else 1.0) * (2.0 * j) ** -gamma * _harmonic_power(j, x)
               for j in range(1, jmax + 1))
    got = (oracle.ssrg_shape(x, kappa, beta, gamma, zeta) -
           oracle.ssrg_shape(x, kappa, 0.0, gamma, zeta)) / beta
    np.testing.assert_allclose(got, np.pi ** 2 / 4.0 * want, rtol=1e-8)


def _ice_index(T=260.0, f_ghz=89.0):
    return np.sqrt(oracle.eps_ice(T, f_ghz))


def test_rayleigh_limit_cross_sections_and_series():
    """Particles far below the wavelength: shape factor 1, so the phase
    function is the Rayleigh one, c (1 + cos^2)/2 with c = 9 k^4 |K|^2 V^2 /
    (4 pi) the radar cross section; absorption 3 k V Im K.  The scattering
    cross section is the reference's step sum over the angle,
    c (pi/2) S / 360 with S = sum over theta_i = i pi/359, i = 1..359, of
    (1 + cos^2)/2 sin (the exact value 2c/3 times 359/360, less the sum's
    own error); the Legendre series is the Rayleigh phase matrix's scaled by
    (2/3) / ((pi/2) S / 360): P11 = [1, 0, 1/2], P12 = [-1/2, 0, 1/2],
    P33 = [0, 3/2, 0]."""
    f = 89e9
    k = 2 * np.pi * f / C_LIGHT
    m = _ice_index()
    d = np.array([1e-7, 1.5e-7, 2e-7])
    dd = np.full(3, 1e-7)
    nd = np.array([1e12, 2e12, 5e11])
    mass = 917.0 * np.pi / 6.0 * d ** 3 * 0.3
    st, ext, alb, back, leg, nl = oracle.ssrg(f, d, dd, nd, mass, 0.6, 0.0, m, 0.25, 1.0, 1.66,
                                              0.04)
    assert st == 0
    eps = m * m
    K = (eps - 1) / (eps + 2)
    V = mass / 917.0
    w = nd * dd
    th = np.arange(1, 360) * np.pi / 359
    step = 0.5 * np.pi * np.sum((1 + np.cos(th) ** 2) / 2 * np.sin(th)) / 360
    cback = np.sum(9 * k ** 4 * abs(K) ** 2 * V ** 2 / (4 * np.pi) * w)
    cabs = np.sum(3 * k * V * K.imag * w)
    np.testing.assert_allclose(alb * ext, cback * step, rtol=1e-8)
    np.testing.assert_allclose(ext, cback * step + cabs, rtol=1e-8)
    np.testing.assert_allclose(back, cback, rtol=1e-8)
    np.testing.assert_allclose(step, 2 / 3 * 359 / 360, rtol=2e-5)
    r = 2 / 3 / step
    np.testing.assert_allclose(leg[0, :3], r * np.array([1.0, 0.0, 0.5]), atol=1e-7)
    np.testing.assert_allclose(leg[1, :3], r * np.array([-0.5, 0.0, 0.5]), atol=1e-7)
    np.testing.assert_allclose(leg[2, :3], r * np.array([0.0, 1.5, 0.0]), atol=1e-7)
    assert nl == 2


def test_series_reconstructs_the_phase_function():
    """Aggregates of a few millimetres at 89 GHz: the Legendre series of P11
    evaluated between the Gauss nodes equals the normalised phase function
    lambda^2/(pi C_sca) sum_bins n dD (|S11|^2 + |S22|^2)/2 built from the
    shape factor, |S22| = 3 k^3 |K| V sqrt(shape)/(4 pi), S11 = S22 cos(theta)."""
    f = 89e9
    lam = C_LIGHT / f
    k = 2 * np.pi / lam
    m = _ice_index()
    d = np.linspace(2e-4, 3e-3, 12)
    dd = np.full(12, d[1] - d[0])
    nd = 1e5 * np.exp(-d / 8e-4)
    mass = 0.0121 * d ** 1.9
    kappa, beta, gamma, zeta = 0.19, 0.23, 1.67, 1.0
    st, ext, alb, back, leg, nl = oracle.ssrg(f, d, dd, nd, mass, 0.6, 0.0, m, kappa, beta, gamma,
                                              zeta)
    assert st == 0 and nl > 4
    eps = m * m
    K2 = abs((eps - 1) / (eps + 2)) ** 2
    mu = np.array([-0.93, -0.41, 0.05, 0.52, 0.87])
    th = np.arccos(mu)
    p11 = np.zeros_like(mu)
    for di, ndi, ddi, mi in zip(d, nd, dd, mass):
        shape = np.array([oracle.ssrg_shape(k * di * 0.6 * np.sin(t / 2), kappa, beta, gamma,
                                            zeta) for t in th])
        a22 = (3 * k ** 3 * mi / 917.0 / (4 * np.pi)) ** 2 * K2 * shape
        p11 += 0.5 * (a22 * np.cos(th) ** 2 + a22) * ndi * ddi
    p11 *= lam ** 2 / (np.pi * alb * ext)
    series = np.polynomial.legendre.legval(mu, leg[0, :nl + 1])
    np.testing.assert_allclose(series, p11, rtol=2e-3)


def test_refused_inputs_in_the_oracle():
    """The assertions of rayleigh_gans.f90:590-616: canting other than 0 or
    90 degrees and a non-positive as_ratio are fatal."""
    m = _ice_index()
    args = (89e9, np.array([1e-3]), np.array([1e-4]), np.array([1e4]), np.array([1e-7]))
    assert oracle.ssrg(*args, 0.6, 45.0, m, 0.25, 1.0, 1.66, 0.04)[0] != 0
    assert oracle.ssrg(*args, 0.0, 0.0, m, 0.25, 1.0, 1.66, 0.04)[0] != 0
    assert oracle.ssrg(*args, 0.6, 90.0, m, 0.25, 1.0, 1.66, 0.04)[0] == 0


def test_parameters_from_the_scattering_name():
    assert _capi.scattering_model("mie-sphere")[0] == "mie-sphere"
    name, rg = _capi.scattering_model("ssrg-rt3")
    assert name == "ssrg-rt3" and rg == (0.25, 1.0, 1.66, float(np.float32(0.04)))
    assert _capi.scattering_model("ssrg-rt3_0.19_0.23_1.67")[1] == (0.19, 0.23, 1.67, 1.0)
    assert _capi.scattering_model("ssrg-rt3_0.19_0.23_1.67_0.5")[1] == (0.19, 0.23, 1.67, 0.5)
    for bad in ("ssrg-rt3_0.19", "ssrg-rt3_1_2_3_4_5", "ssrg-rt3_a_b_c"):
        with pytest.raises(ValueError):
            _capi.scattering_model(bad)
    assert _capi.scattering_model("tmatrix")[0] == "tmatrix"
    with pytest.raises(NotImplementedError):
        _capi.scattering_model("ssrga")


# ------------------------------------------------------------ device parity


SCATS = ["ssrg-rt3", "ssrg-rt3_0.19_0.23_1.67", "ssrg-rt3_0.19_0.23_1.67_0.5"]
FREQS = [10.65, 35.5, 89.0, 183.31]


def _rain_and_aggregates(scat, ar=0.6, cant=-99.0, liq_ice=-1):
    pam = pamtra_amd.pyPamtra()
    pam.df.addHydrometeor(("rwc_q", -99.0, 1, -99.0, -99.0, -99.0, -99.0, -99.0, 3, 30, "exp",
                           -99.0, -99.0, 8.0e6, -99.0, 1.0e-5, 6.0e-3, "mie-sphere",
                           "khvorostyanov01_drops", -99.0))
    pam.df.addHydrometeor(("swc_q", ar, liq_ice, -99.0, 0.038, 2.0, 0.3971, 1.88, 3, 50,
                           "exp_field_t", -99.0, -99.0, -99.0, -99.0, 0.51e-10, 2.0e-2, scat,
                           "heymsfield10_particles", cant))
    pam = importer.createUsStandardProfile(pam, hgt_lev=np.array([0.0, 1000.0, 2000.0, 3000.0]))
    pam.p["hydro_q"][0, 0, 1, 0] = 1e-3
    pam.p["hydro_q"][0, 0, 1:3, 1] = 5e-4
    return pam


def _aggregate_columns(ncol, seed):
    """The four-category columns with snow and graupel as aggregates."""
    pam, freqs = workloads.config3(ncol=ncol, seed=seed)
    pam.df.data["scat_name"][2] = "ssrg-rt3_0.19_0.23_1.67"
    pam.df.data["as_ratio"][2] = 0.6
    pam.df.data["scat_name"][3] = "ssrg-rt3"
    pam.df.data["as_ratio"][3] = 0.8
    pam.df.data["canting"][3] = 90.0
    return pam, freqs


def _gpu(pam, freqs, nml=None, precision="fp64"):
    from pamtra_amd import _lib
    lib = _lib.library()
    ctx = _lib.get_context(0)
    m = _capi.marshal(nml or passive_nml(pam), dict(pam.set, precision=precision), pam.df.data,
                      pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    print(f"pam_run ssrg ncol={m.nx * m.ny} L={m.max_nlyrs} nf={len(f)} {precision}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    out = dict(hr.arrays)
    out["msg"] = _capi.results_message(hr)
    return out, st


def _cpu(pam, freqs, nml=None):
    out, st, _ = oracle.run(nml or passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p,
                            freqs, nthreads=8, diagnostics=True)
    return out, st


def _compare(g, c, tol=1e-3):
    ok = np.isfinite(c["tb"])
    assert ok.any() and np.array_equal(ok, np.isfinite(g["tb"]))
    d = np.abs(g["tb"][ok] - c["tb"][ok]).max()
    print(f"max |dTB| = {d:.3e} K", flush=True)
    assert d <= tol
    np.testing.assert_array_equal(g["nlegen"], c["nlegen"])
    np.testing.assert_array_equal(g["ndoubles"], c["ndoubles"])
    m = np.isfinite(c["kexttot"])
    np.testing.assert_allclose(g["kexttot"][m], c["kexttot"][m], rtol=1e-9, atol=1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("scat", SCATS)
def test_aggregates_with_rain_vs_oracle(scat):
    pam = _rain_and_aggregates(scat)
    g, st = _gpu(pam, FREQS)
    c, st2 = _cpu(pam, FREQS)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    _compare(g, c)


@pytest.mark.gpu
def test_aggregate_columns_vs_oracle_and_mixed_precision():
    pam, freqs = _aggregate_columns(24, 2101)
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    _compare(g, c)
    mx, st3 = _gpu(pam, freqs, precision="mixed")
    assert st3 == 0
    d = np.nanmax(np.abs(mx["tb"] - g["tb"]))
    print(f"mixed vs fp64 max |dTB| = {d:.3e} K", flush=True)
    assert d <= 0.05


@pytest.mark.gpu
def test_aggregate_reflectivity_vs_oracle():
    pam, freqs = _aggregate_columns(12, 2102)
    nml = dict(pam.nmlSet)
    nml.update(active=True, radar_attenuation="top-down", radar_polarisation="NN,VV,HV")
    g, st = _gpu(pam, freqs, nml)
    c, st2 = _cpu(pam, freqs, nml)
    assert st == st2 == 0
    assert np.array_equal(g["ze"] == -9999.0, c["ze"] == -9999.0)
    ok = c["ze"] > -9999.0
    assert ok.any()
    d = np.max(np.abs(g["ze"][ok] - c["ze"][ok]))
    print(f"max |dZe| = {d:.2e} dB", flush=True)
    assert d <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["liquid", "canting", "as_ratio"])
def test_refused_inputs_on_the_device(case):
    kw = {"liquid": dict(liq_ice=1), "canting": dict(cant=45.0), "as_ratio": dict(ar=-99.0)}[case]
    pam = _rain_and_aggregates("ssrg-rt3", **kw)
    g, st = _gpu(pam, FREQS)
    c, st2 = _cpu(pam, FREQS)
    assert st != 0 and st2 != 0
    assert "rayleigh_gans" in g["msg"], g["msg"]


@pytest.mark.gpu
def test_aggregates_at_32_angles_vs_oracle():
    """The 32-angle library (64x64 operators) with the aggregate producer,
    against the 32-angle oracle."""
    from pamtra_amd import _lib
    pam = _rain_and_aggregates("ssrg-rt3_0.19_0.23_1.67")
    sets = dict(pam.set, _nangles=32)
    m = _capi.marshal(passive_nml(pam), sets, pam.df.data, pam.df.data4D, pam.p)
    assert m.nummu == 32
    lib, ctx = _lib.library(32), _lib.get_context(0, 32)
    f, fp = _capi.freqs_array(FREQS)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    print("pam_run ssrg 32 angles", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    g = dict(hr.arrays)
    c, st2, _ = oracle.run(passive_nml(pam), sets, pam.df.data, pam.df.data4D, pam.p, FREQS,
                           nthreads=8, diagnostics=True)
    assert st == 0 and st2 == 0
    assert g["tb"].shape[3] == 64
    _compare(g, c)


def test_other_scattering_names_are_refused_with_the_reason():
    for name, word in (("ssrga_0.2", "unset"), ("ss-rayleigh-gans", "asserts"),
                       ("rayleigh", "backscatter-only"), ("rayleigh-gans", "backscatter-only"),
                       ("liudb_8", "DDA-database"),
                       ("hongdb_3", "DDA-database"), ("nonsense", "unknown")):
        with pytest.raises(NotImplementedError, match=word):
            _capi.scattering_model(name)
