"""Pins the CPU oracle (oracle/pam_oracle.c) before it is used to check the
device path: the reference's own golden vectors and known answers, closed
forms, an independent Mie series, invariants and a brute-force matrix
exponential."""

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg
from scipy.linalg import expm

import oracle
from pamtra_amd import workloads
from _scenario import FREQUENCIES, build_pamtra, passive_nml

GOLDEN = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__),
                                            "golden", "rain_snow_reference.npz"))


def _run(pam, freqs, **kw):
    out, st, _ = oracle.run(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p, freqs,
                            **kw)
    return out, st


def test_golden_brightness_temperature():
    """reference tests/test_regression.py:42-43 (rtol 1e-5)."""
    out, st = _run(build_pamtra(), FREQUENCIES)
    assert st == 0
    np.testing.assert_allclose(out["tb"], GOLDEN["tb"], rtol=1e-5)
    assert np.max(np.abs(out["tb"] - GOLDEN["tb"])) < 1e-4


def test_golden_is_sensitive_to_the_profile():
    """Negative control: the geometric-height standard atmosphere moves TB by
    ~8 mK, which the golden comparison must reject."""
    out, _ = _run(build_pamtra(geopotential=False), FREQUENCIES)
    assert np.max(np.abs(out["tb"] - GOLDEN["tb"])) > 1e-3


def test_tkc_known_answer():
    """reference tests/test_fortran_functions.py:23-33."""
    e = oracle.eps_water_tkc(20.0, 10.0)
    np.testing.assert_allclose([e.real, e.imag], [60.6084, 32.9184], rtol=1e-3)


@pytest.mark.parametrize("nummu", [16, 32])
def test_lobatto_nodes_are_roots_of_legendre_derivative(nummu):
    mu, w = oracle.lobatto(nummu)
    n = 2 * nummu
    # interior Lobatto nodes of the 2N-point rule: roots of P'_{n-1}
    c = np.zeros(n)
    c[-1] = 1.0
    roots = np.sort(npleg.legroots(npleg.legder(c)))
    pos = roots[roots > 0]
    np.testing.assert_allclose(mu[:-1], pos, rtol=0, atol=1e-13)
    assert mu[-1] == 1.0
    pn1 = npleg.legval(mu, c)
    np.testing.assert_allclose(w, 2.0 / (n * (n - 1) * pn1 ** 2), rtol=1e-11)
    # the full rule integrates x^2 exactly over [-1, 1] (2/3)
    assert abs(2 * np.sum(w * mu ** 2) - 2.0 / 3.0) < 1e-12


@pytest.mark.parametrize("n", [5, 29, 45, 77, 121])
def test_gauss_legendre_nodes(n):
    x, w = oracle.gausquad(n)
    xr, wr = npleg.leggauss(n)
    np.testing.assert_allclose(x, xr, atol=1e-13)
    np.testing.assert_allclose(w, wr, rtol=1e-9)


def _bhmie(x, m, nstop):
    """Independent Mie efficiencies (Bohren & Huffman 1983, appendix A),
    full double precision, complex xi recursion.  The downward recursion
    starts at nstop + 15 as in the reference (mie_scat_utilities.f90:46),
    not at max(nstop, |mx|) + 15 as Bohren & Huffman recommend."""
    mx = m * x
    nmx = int(nstop + 15)
    d = np.zeros(nmx + 1, dtype=complex)
    for n in range(nmx, 0, -1):
        d[n - 1] = n / mx - 1.0 / (d[n] + n / mx)
    psi0, psi1 = np.cos(x), np.sin(x)
    chi0, chi1 = -np.sin(x), np.cos(x)
    xi1 = complex(psi1, -chi1)
    qext = qsca = 0.0
    for n in range(1, nstop + 1):
        psi = (2 * n - 1) * psi1 / x - psi0
        chi = (2 * n - 1) * chi1 / x - chi0
        xi = complex(psi, -chi)
        an = ((d[n] / m + n / x) * psi - psi1) / ((d[n] / m + n / x) * xi - xi1)
        bn = ((m * d[n] + n / x) * psi - psi1) / ((m * d[n] + n / x) * xi - xi1)
        qext += (2 * n + 1) * (an.real + bn.real)
        qsca += (2 * n + 1) * (abs(an) ** 2 + abs(bn) ** 2)
        psi0, psi1, chi0, chi1, xi1 = psi1, psi, chi1, chi, xi
    return 2 * qext / x ** 2, 2 * qsca / x ** 2


@pytest.mark.parametrize("x,m", [(0.05, 7.0 - 2.5j), (0.8, 5.9 - 2.9j), (3.5, 3.2 - 1.8j),
                                 (12.0, 4.5 - 2.4j), (25.0, 1.78 - 0.003j), (40.0, 1.3 - 0.0j)])
def test_mie_against_independent_series(x, m):
    nt, qe, qs, qb = oracle.mie_single(x, m)
    # the oracle's conj-convention index has a negative imaginary part
    qe2, qs2 = _bhmie(x, m.conjugate(), nt)
    # xi passes through single precision in the reference (miecalc)
    np.testing.assert_allclose([qe, qs], [qe2, qs2], rtol=2e-6)


def test_mie_real_index_conserves_energy():
    nt, qe, qs, qb = oracle.mie_single(6.0, 1.33 + 0j)
    assert abs(qe - qs) / qe < 1e-6


def test_mie_rayleigh_limit():
    x, m = 1e-3, 7.0 - 2.5j
    nt, qe, qs, qb = oracle.mie_single(x, m)
    K = (m ** 2 - 1) / (m ** 2 + 2)
    np.testing.assert_allclose(qs, 8.0 / 3.0 * x ** 4 * abs(K) ** 2, rtol=1e-3)
    np.testing.assert_allclose(qe, 4 * x * (-K.imag) + qs, rtol=1e-3)


def _hg_coefs(nlegen, g):
    l = np.arange(nlegen + 1)
    c1 = (2 * l + 1) * g ** l
    c2 = -0.2 * c1 * (l >= 2)
    c3 = 0.9 * c1
    return np.stack([c1, c2, c3])


def test_scatcnv_normalisation_and_symmetry():
    coef = _hg_coefs(24, 0.6)
    kext, salb = 2e-3, 0.8
    smat = oracle.scatcnv(24, coef, kext, salb)
    mu, w = oracle.lobatto(16)
    # quadrant symmetry: (-,-) = (+,+), (+,-) = (-,+)
    np.testing.assert_array_equal(smat[3], smat[0])
    np.testing.assert_array_equal(smat[2], smat[1])
    # CHECK_NORM4: emission - extinction + 2 pi sum w (P++ + P-+) ~ 0
    for j1 in range(16):
        s = (1 - salb) * kext - kext + 2 * np.pi * np.sum(
            w * (smat[0][0::2, 2 * j1] + smat[2][0::2, 2 * j1]))
        assert abs(s) < 5e-4 * kext


def _exact_layer(smat, kext, h):
    """R and T of a homogeneous source-free layer from the exponential of
    the two-stream-set generator (interaction principle)."""
    n = smat.shape[-1]
    mu, w = oracle.lobatto(n // 2)
    muv = np.repeat(mu, 2)
    wv = np.repeat(w, 2)
    Minv = np.diag(1 / muv)
    K = np.eye(n) * kext
    A = Minv @ (K - 2 * np.pi * smat[0] * wv[None, :])
    B = Minv @ (2 * np.pi * smat[1] * wv[None, :])
    G = np.block([[-A, B], [-B, A]])
    Phi = expm(h * G)
    P21, P22 = Phi[n:, :n], Phi[n:, n:]
    P22i = np.linalg.inv(P22)
    return -P22i @ P21, P22i


# At 32 angles the smallest node is mu = 0.0126, so exp(tau / mu) of the
# generator exceeds the double range of the brute force beyond tau ~ 0.5.
@pytest.mark.parametrize("tau,salb,nummu", [(0.05, 0.5, 16), (0.4, 0.9, 16), (1.2, 0.3, 16),
                                            (0.05, 0.5, 32), (0.4, 0.9, 32), (0.3, 0.3, 32)])
def test_doubling_matches_matrix_exponential(tau, salb, nummu):
    coef = _hg_coefs(16, 0.5)
    kext = 1e-3
    smat = oracle.scatcnv(16, coef, kext, salb, nummu)
    h = tau / kext
    st, R, T, sp, sm, nd = oracle.layer_doubling(smat, kext, salb, 0.0, h, 0.0, 0.0, 3000.0)
    assert st == 0 and nd > 10
    Rx, Tx = _exact_layer(smat, kext, h)
    assert np.max(np.abs(R - Rx)) < 2e-5 * max(1.0, np.max(np.abs(Rx)))
    assert np.max(np.abs(T - Tx)) < 2e-5


def test_isothermal_layer_source_is_kirchhoff():
    """An isothermal layer emits (I - R - T) B on the intensity components."""
    coef = _hg_coefs(12, 0.4)
    kext, salb, T0, wl = 2e-3, 0.6, 270.0, 3370.0
    smat = oracle.scatcnv(12, coef, kext, salb)
    st, R, T, sp, sm, nd = oracle.layer_doubling(smat, kext, salb, 1e-5, 400.0, T0, T0, wl)
    B = oracle.planck(T0, wl)
    b = np.zeros(32)
    b[0::2] = B
    expect = (np.eye(32) - R - T) @ b
    # intensity components: exact up to the thin-layer start (radtran4.f90:
    # 323-329, 1e-6 optical depth); the Q emission is zero in the reference
    # (scatcnv.f90:90-98), so Q is not Kirchhoff-consistent and not checked
    np.testing.assert_allclose(sp[0::2], expect[0::2], rtol=2e-4)
    np.testing.assert_allclose(sm[0::2], expect[0::2], rtol=2e-4)


def test_planck_inverse():
    for t in (2.73, 150.0, 288.15, 320.0):
        for wl in (923.0, 3370.0, 28150.0):
            assert abs(oracle.rad2tb(oracle.planck(t, wl), wl) - t) < 1e-9 * t


def test_isothermal_black_column():
    """Isothermal atmosphere over a black surface at the same temperature:
    without hydrometeors the upwelling TB equals that temperature at every
    level and angle; with a scattering layer the layer reflects the 2.73 K
    sky, so the top-of-atmosphere TB drops below it while the surface-level
    upwelling (the black surface alone) stays exact."""
    T0 = 275.0
    for hq in (0.0, 2e-3):
        pam = build_pamtra()
        pam.p["temp_lev"][:] = T0
        pam.p["groundtemp"] = np.full((1, 1), T0)
        pam.nmlSet["emissivity"] = 1.0
        pam.p["hydro_q"][:] = 0.0
        pam.p["hydro_q"][0, 0, 1, 0] = hq
        out, st = _run(pam, [35.5, 89.0])
        assert st == 0
        up = out["tb"][0, 0, :, :16, :, :]
        if hq == 0.0:
            np.testing.assert_allclose(up, T0, atol=1e-9)
        else:
            np.testing.assert_allclose(up[1], T0, atol=1e-9)
            assert np.all(up[0] < T0) and np.all(up[0] > 100.0)


def test_gas_absorption_dry_air_positive_and_monotone_in_pressure():
    """Dry-air absorption grows with pressure (pressure-broadened O2 wings)
    and water vapour only adds absorption.  H2O line strengths themselves are
    parity unpinned (no reference fixture covers humid air)."""
    a1 = oracle.gasabs(35.5, 280.0, 0.0, 50000.0)
    a2 = oracle.gasabs(35.5, 280.0, 0.0, 100000.0)
    a3 = oracle.gasabs(35.5, 280.0, 0.01, 100000.0)
    assert 0 < a1 < a2 < a3
    assert oracle.gasabs(900.0, 280.0, 0.0, 1e5) < 0  # invalid input is refused


def _liquid_row(dist, nbin, d1, d2, p1=-99.0, p2=-99.0, p3=-99.0, p4=-99.0):
    return {"hydro_name": "x", "as_ratio": -99.0, "liq_ice": 1, "rho_ms": -99.0,
            "a_ms": -99.0, "b_ms": -99.0, "alpha_as": -99.0, "beta_as": -99.0,
            "moment_in": 3, "nbin": nbin, "dist_name": dist, "p_1": p1, "p_2": p2,
            "p_3": p3, "p_4": p4, "d_1": d1, "d_2": d2, "scat_name": "mie-sphere",
            "vel_size_mod": "khvorostyanov01_drops", "canting": -99.0}


@pytest.mark.parametrize("case", ["mgamma", "exp"])
def test_dsd_parameters_reproduce_the_prescribed_moments(case):
    """Closed form of make_dist_params (make_dist_param.f90): without the
    mass rescale, the binned spectrum of a modified-gamma cloud (N_tot and q
    given, mu=2, gamma=1) and of a Marshall-Palmer rain (N0 and q given)
    integrates to the prescribed number and mass, for liquid spheres
    m = pi/6 rho_w D^3.  A wrong exponent or Gamma argument in the parameter
    derivation misses these by a large factor."""
    nml = {"conserve_mass_rescale_dsd": False, "hydro_adaptive_grid": False}
    a, b = np.pi / 6.0 * 1000.0, 3.0
    if case == "mgamma":
        q, ntot = 2.0e-4, 1.0e8
        row = _liquid_row("mgamma", 4000, 1.0e-7, 2.0e-4, p1=ntot, p3=2.0, p4=1.0)
    else:
        q, n0 = 5.0e-4, 8.0e6
        row = _liquid_row("exp", 4000, 1.0e-6, 2.0e-2, p3=n0)
    st, (d, dd, n, rho, dscat) = oracle.dsd(row, nml, q, -99.0, -99.0, 280.0)
    assert st == 0
    mass = np.sum(a * d ** b * n)
    np.testing.assert_allclose(mass, q, rtol=2e-3)
    if case == "mgamma":
        np.testing.assert_allclose(np.sum(n), ntot, rtol=2e-3)
    else:
        # N(D) = N0 exp(-lambda D) with lambda = (a N0 Gamma(b+1) / q)^(1/(b+1))
        lam = (a * n0 * 6.0 / q) ** 0.25
        keep = n > 1e-3 * n.max()
        np.testing.assert_allclose(n[keep] / dd[keep], n0 * np.exp(-lam * d[keep]), rtol=2e-3)


def test_goff_gratch_known_answers():
    """e_sat_gg_water: the reference's known answer at 0 C
    (tests/test_meteosi.py:6-10, 610.34 Pa) and the steam point the
    Goff-Gratch formula is anchored at (1013.246 hPa at 373.16 K)."""
    np.testing.assert_allclose(oracle.esat(273.15) * 100.0, 610.34, rtol=1e-3)
    np.testing.assert_allclose(oracle.esat(373.16), 1013.246, rtol=1e-6)


def test_water_vapour_line_peaks_at_22_235_ghz():
    """abh2o: at 100 hPa, where the line is narrow, humid-air absorption
    peaks at the 22.235 GHz line centre; absorption grows with vapour
    density (the line table itself is not pinned)."""
    f = np.linspace(20.0, 24.5, 451)
    k = np.array([oracle.gasabs(x, 293.0, 7.5e-3, 10000.0) for x in f])
    assert abs(f[np.argmax(k)] - 22.235) <= 0.02
    assert oracle.gasabs(22.235, 293.0, 15e-3, 101325.0) > 1.5 * oracle.gasabs(
        22.235, 293.0, 7.5e-3, 101325.0)


def test_tessem2_bounds_and_polarisation_order():
    """TESSEM2 ocean emissivity: within (0, 1), vertical above horizontal at
    oblique angles, equal at nadir within the network's accuracy."""
    for f in (10.65, 37.0, 89.0, 183.31):
        for zen in (0.0, 30.0, 53.0, 70.0):
            (ev, eh), _ = oracle.tessem2(f, zen, 5.0, 290.0, 35.0)
            assert 0.0 < eh < 1.0 and 0.0 < ev < 1.0
            if zen >= 30.0:
                assert ev > eh
        (ev0, eh0), _ = oracle.tessem2(f, 0.0, 5.0, 290.0, 35.0)
        assert abs(ev0 - eh0) < 0.02


def _run_active(pam, freqs, **nml_over):
    nml = dict(pam.nmlSet)
    nml["active"] = True
    nml.update(nml_over)
    out, st, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, freqs, nthreads=4)
    return out, st


def test_golden_radar_reflectivity_and_attenuation():
    """reference tests/test_regression.py:36-40,46-49: the radar simulator's
    "simple" mode (the default active run) on the golden scenario."""
    out, st = _run_active(build_pamtra(), FREQUENCIES)
    assert st == 0
    np.testing.assert_allclose(out["ze"], GOLDEN["Ze"], rtol=1e-5)
    np.testing.assert_allclose(out["att_hydro"], GOLDEN["Att_hydro"], rtol=1e-5)
    assert (out["ze"][0, 0, 1:3] > -9999).all()
    assert (out["ze"][0, 0, 0] == -9999).all()  # no hydrometeors in the lowest layer


def test_radar_attenuation_modes_and_polarisations():
    """The path-integrated attenuation of radar_simulator.f90:130-150 as a
    difference of reflectivities, from the attenuation outputs: bottom-up
    2 sum_{k<z}(Ah+Aa) + Ah+Aa at z; top-down 2 sum_{k>z} Aa + Ah+Aa at z
    (the layers above are visited later, their Ah still 0).  Cross-polar
    channels of spheres have no backscattering."""
    pam, freqs = workloads.config2(ncol=4, seed=61)
    base, st = _run_active(pam, freqs, radar_polarisation="NN,HV")
    assert st == 0
    ze0 = base["ze"][..., 0, 0]
    assert np.all(base["ze"][..., 1, 0] == -9999.0)
    ah, aa = base["att_hydro"][..., 0], base["att_atmo"]
    has = ze0 > -9999.0
    assert has.any()
    bu, _ = _run_active(pam, freqs, radar_attenuation="bottom-up")
    td, _ = _run_active(pam, freqs, radar_attenuation="top-down")
    nl = pam.p["nlyrs"]
    for ix in range(4):
        L = int(nl[ix, 0])
        for f in range(len(freqs)):
            for z in range(L):
                if not has[ix, 0, z, f]:
                    continue
                pia_bu = 2 * (ah[ix, 0, :z, f].sum() + aa[ix, 0, :z, f].sum()) + ah[ix, 0, z, f] + aa[ix, 0, z, f]
                pia_td = 2 * aa[ix, 0, z + 1:L, f].sum() + ah[ix, 0, z, f] + aa[ix, 0, z, f]
                np.testing.assert_allclose(ze0[ix, 0, z, f] - bu["ze"][ix, 0, z, f, 0, 0], pia_bu,
                                           rtol=1e-9, atol=1e-9)
                np.testing.assert_allclose(ze0[ix, 0, z, f] - td["ze"][ix, 0, z, f, 0, 0], pia_td,
                                           rtol=1e-9, atol=1e-9)
    # one-way attenuation of the gas: 10 log10(exp(kgas dz))
    hl = pam.p["hgt_lev"]
    dz = hl[:, :, 1:] - hl[:, :, :-1]
    diag, _, _ = oracle.run(dict(pam.nmlSet, active=False), pam.set, pam.df.data, pam.df.data4D,
                            pam.p, freqs, nthreads=4, diagnostics=True)
    kg = diag["kextatmo"]
    for ix in range(4):
        L = int(nl[ix, 0])
        np.testing.assert_allclose(aa[ix, 0, :L], 10 * np.log10(np.exp(kg[ix, 0, :L] * dz[ix, 0, :L, None])),
                                   rtol=1e-12, atol=1e-14)
        assert np.all(aa[ix, 0, L:] == 0.0) and np.all(base["radar_hgt"][ix, 0, L:] == -9999.0)


def _planck(t, wl):
    """Planck radiance [W m-2 sr-1 um-1] at wavelength wl [um] with the
    constants of rt_utilities.f90:69-90."""
    return 1.1911e8 / wl ** 5 / (np.exp(1.4388e4 / (wl * t)) - 1.0)


def _crossing(i0, taus, b_in, b_out, mu):
    """Radiance after crossing clear layers in order, each with the Planck
    radiance linear in optical depth from b_in (entry side) to b_out: the
    solution of mu dI/dtau = B - I by adaptive quadrature."""
    from scipy.integrate import quad
    i = i0
    for tau, b0, b1 in zip(taus, b_in, b_out):
        if tau <= 0.0:
            continue
        src = quad(lambda x: (b0 + (b1 - b0) * x / tau) * np.exp(-(tau - x) / mu) / mu, 0.0, tau,
                   epsabs=0.0, epsrel=1e-13, limit=200)[0]
        i = i * np.exp(-tau / mu) + src
    return i


def _with_levels(pam, heights):
    nx, ny = pam.p["ngridx"], pam.p["ngridy"]
    pam.p["noutlevels"] = len(heights[0])
    pam.p["obs_height"] = np.asarray(heights, dtype=float).reshape(nx, ny, len(heights[0]))
    return pam


@pytest.mark.parametrize("nummu", [16, 32])
def test_interior_output_levels_clear_column_closed_form(nummu):
    """Output levels inside the column (radtran4.f90:375-422, run_rt.f90:166-184):
    a clear, gas-absorbing column over a specular surface of emissivity 0.7.
    Up- and downwelling TB at every level against the radiative transfer
    equation integrated by quadrature, with the level nearest to each
    observation height (ties to the upper level)."""
    pam, freqs = workloads.config1(ncol=2, seed=81)
    pam.p["hydro_q"][:] = 0.0
    pam.p["sfc_type"][:] = -9999
    pam.p["sfc_refl"][:] = "S"
    pam.nmlSet["emissivity"] = 0.7
    hl = pam.p["hgt_lev"][:, 0, :]
    dzl = hl[:, 1:] - hl[:, :-1]
    heights, expect_lev = [], []
    for c in range(2):
        h = hl[c]
        obs = [833000.0, 0.0, h[10], h[5] + 0.3 * dzl[c, 5], h[20] + 0.8 * dzl[c, 20],
               h[15] + 0.5 * dzl[c, 15], 1e9, -5.0]
        heights.append(obs)
        expect_lev.append([30, 0, 10, 5, 21, 16, 30, 0])
    pam = _with_levels(pam, heights)
    out, st, _ = oracle.run(passive_nml(pam), dict(pam.set, _nangles=nummu), pam.df.data,
                            pam.df.data4D, pam.p, freqs, nthreads=2, diagnostics=True)
    assert st == 0
    mu, _ = oracle.lobatto(nummu)
    wl = 299792458.0 / (freqs[0] * 1e3)
    for c in range(2):
        t_lev = pam.p["temp_lev"][c, 0]
        L = int(pam.p["nlyrs"][c, 0])
        tau = dzl[c, :L] * np.maximum(out["kextatmo"][c, 0, :L, 0], 0.0)
        assert abs(_planck(t_lev[3], wl) - oracle.planck(t_lev[3], wl)) < 1e-13 * _planck(t_lev[3], wl)
        bl = _planck(t_lev[:L + 1], wl)
        bg = _planck(pam.p["groundtemp"][c, 0], wl)
        for j in range(nummu):
            m = mu[j]
            down = [_planck(2.73, wl)]  # downwelling at level L, L-1, ..., 0
            for z in range(L - 1, -1, -1):
                down.append(_crossing(down[-1], [tau[z]], [bl[z + 1]], [bl[z]], m))
            down = down[::-1]  # index = level
            up = [0.7 * bg + 0.3 * down[0]]
            for z in range(L):
                up.append(_crossing(up[-1], [tau[z]], [bl[z]], [bl[z + 1]], m))
            for io, k in enumerate(expect_lev[c]):
                tb_up = oracle.rad2tb(up[k], wl)
                tb_dn = oracle.rad2tb(down[k], wl)
                got = out["tb"][c, 0, io, :, 0, :]
                np.testing.assert_allclose(got[nummu - 1 - j], tb_up, atol=2e-7, rtol=0)
                np.testing.assert_allclose(got[nummu + j], tb_dn, atol=2e-7, rtol=0)


def test_interior_level_above_transparent_layers_sees_the_surface_level():
    """With no gas extinction the layers below a level are the identity, so
    with hydrometeors only above it the level's radiances equal the surface
    level's: the joint of the two stacks (both gammas non-trivial) at the
    interior level against the surface-alone joint."""
    pam, freqs = workloads.config1(ncol=3, seed=82)
    hl = pam.p["hgt_lev"][:, 0, :]
    pam.p["hydro_q"][:, :, :12, :] = 0.0
    pam.p["hydro_q"][:, :, 12:18, 0] = 3e-4
    pam.nmlSet["lgas_extinction"] = False
    pam = _with_levels(pam, [[833000.0, 0.0, hl[c, 12], hl[c, 6]] for c in range(3)])
    out, st = _run(pam, freqs, nthreads=3)
    assert st == 0
    tb = out["tb"][:, 0]
    np.testing.assert_allclose(tb[:, 2], tb[:, 1], rtol=0, atol=1e-9)
    np.testing.assert_allclose(tb[:, 3], tb[:, 1], rtol=0, atol=1e-9)
    assert np.all(np.abs(tb[:, 0, :16] - tb[:, 1, :16]) > 1e-3)  # scattering above the level


def test_thirty_two_angles_converge_at_nadir():
    """The 32-angle build (nummu = 32, as a recompiled reference would run)
    on the golden scenario: every TB at 64 angles is finite, the two nadir
    TB (mu = 1, up at the top, down at the surface) agree with the 16-angle
    build's within the quadrature's convergence, and an isothermal column
    over a black surface at the same temperature gives that temperature at
    all 64 angles."""
    pam = build_pamtra()
    a, st = _run(pam, FREQUENCIES)
    b, st2, _ = oracle.run(passive_nml(pam), dict(pam.set, _nangles=32), pam.df.data,
                           pam.df.data4D, pam.p, FREQUENCIES)
    assert st == st2 == 0 and b["tb"].shape == (1, 1, 2, 64, 2, 2)
    assert np.all(np.isfinite(b["tb"]))
    np.testing.assert_allclose(b["tb"][0, 0, 0, 0], a["tb"][0, 0, 0, 0], atol=0.05)
    np.testing.assert_allclose(b["tb"][0, 0, 1, 63], a["tb"][0, 0, 1, 31], atol=0.05)
    T0 = 265.0
    pam.p["temp_lev"][:] = T0
    pam.p["groundtemp"] = np.full((1, 1), T0)
    pam.nmlSet["emissivity"] = 1.0
    pam.p["hydro_q"][:] = 0.0
    c, st3, _ = oracle.run(passive_nml(pam), dict(pam.set, _nangles=32), pam.df.data,
                           pam.df.data4D, pam.p, [35.5])
    assert st3 == 0
    np.testing.assert_allclose(c["tb"][0, 0, :, :32], T0, atol=1e-9)


def test_liu_sea_water_permittivity_known_answers():
    """FASTEM's sea-water permittivity (liu.f90): static permittivity of pure
    water at 20 C (80.2, e.g. Kaatze 1989) and the ionic conductivity that
    defines practical salinity (PSS-78: S = 35 at 15 C is 4.2914 S/m), read
    from the imaginary part at 10 MHz where conduction dominates."""
    eps0 = 1.0 / (np.pi * 4.0e-7 * 299792458.0 ** 2)
    e = oracle.liu_permittivity(293.15, 0.0, 1e-4)
    assert abs(e.real - 80.2) < 0.1
    e = oracle.liu_permittivity(288.15, 35.0, 0.01)
    sigma = -e.imag * 2 * np.pi * 0.01e9 * eps0
    assert abs(sigma - 4.2914) < 0.005, sigma
    # optical limit: epsilon_inf of the model, a few units
    assert 3.5 < oracle.liu_permittivity(293.15, 35.0, 1e5).real < 5.0


def test_fastem_emissivity_limits_and_fresnel():
    """FASTEM5 (fastemx.f90, with PAMTRA's call: no azimuth or reflection
    correction).  Calm sea at 70 degrees or more: no roughness term and no
    foam, so e = 1 - Fresnel reflectivity of the Liu permittivity, checked
    against Snell's law with a complex refraction angle.  Within the fit's
    frequency range (up to 200 GHz) 0 < e < 1, and V above H off nadir; at
    325 GHz the large-scale term, a quadratic in frequency, can push e_v
    past 1 near 70 degrees, in the reference as here.  Parity unpinned beyond
    this: no reference fixture covers FASTEM."""
    for f in (10.65, 37.0, 89.0, 183.0):
        for zen in (70.0, 75.0, 80.0):
            eps = oracle.liu_permittivity(290.0, 35.0, f)
            th = np.radians(zen)
            n = np.sqrt(eps)
            ct = np.sqrt(1 - (np.sin(th) / n) ** 2)
            rh = abs((np.cos(th) - n * ct) / (np.cos(th) + n * ct)) ** 2
            rv = abs((n * np.cos(th) - ct) / (n * np.cos(th) + ct)) ** 2
            e, r = oracle.fastem(f, zen, 290.0, 35.0, 0.0)
            np.testing.assert_allclose(e, [1 - rv, 1 - rh], rtol=0, atol=2e-12)
            np.testing.assert_allclose(r, 1 - e, rtol=0, atol=1e-15)
    for f in (6.9, 23.8, 50.3, 150.0, 200.0):
        for zen in (0.0, 20.0, 45.0, 60.0, 69.0):
            for wind in (0.0, 5.0, 15.0, 30.0):
                e, _ = oracle.fastem(f, zen, 285.0, 33.0, wind)
                assert np.all((e > 0) & (e < 1)), (f, zen, wind, e)
                if zen >= 20.0:
                    assert e[0] > e[1], (f, zen, wind, e)
    e, _ = oracle.fastem(325.0, 69.0, 285.0, 33.0, 5.0)
    assert e[0] > 1.0  # the extrapolated fit, kept as the reference has it


def test_fastem_and_tessem2_agree_roughly_near_nadir():
    """The two ocean models of sfc_optics.f90:55-58, calm sea: the mean of
    the V and H emissivities within 0.03 of each other up to 30 degrees
    (further off nadir they part by up to 0.06, and near grazing TESSEM2's
    network even returns a negative H emissivity).  A check across two
    independent published models, not a parity pin."""
    for f in (10.65, 37.0, 89.0):
        for sst in (280.0, 290.0, 300.0):
            for zen in (0.0, 15.0, 30.0):
                a, _ = oracle.fastem(f, zen, sst, 33.0, 0.0)
                b, _ = oracle.tessem2(f, zen, 0.0, sst, 33.0)
                assert abs(a.mean() - b.mean()) < 0.03, (f, sst, zen, a, b)
