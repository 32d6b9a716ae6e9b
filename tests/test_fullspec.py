"""Full-spectrum input (nmlSet hydro_fullspec with df.addFullSpectra;
reference hydrometeor_extinction.f90:120-150, scatProperties.f90:258-287):
the given bins replace the size distribution.

The pin: a run whose full spectrum holds the very bins the size
distribution produces (rain as Mie spheres, aggregates as self-similar
Rayleigh-Gans particles with the descriptor's parameters as per-bin lists)
gives the brightness temperatures of the ordinary run, in the oracle and on
the device.  The device then matches the oracle with per-bin parameters that
vary."""

import ctypes as C

import numpy as np
import pytest

import oracle
import pamtra_amd
from pamtra_amd import _capi, importer
from _scenario import passive_nml

FREQS = [10.65, 35.5, 89.0, 183.31]
# moment_in 0: the spectrum comes from the descriptor alone, the same in
# every layer, so its bins can be taken from the size-distribution routine
RAIN = ("rain", -99.0, 1, -99.0, -99.0, -99.0, -99.0, -99.0, 0, 30, "exp", 8e6, 2000.0, -99.0,
        -99.0, 1e-4, 6e-3, "mie-sphere", "khvorostyanov01_drops", -99.0)
SNOW = ("snow", 0.6, -1, -99.0, 0.038, 2.0, -99.0, -99.0, 0, 30, "exp", 3e6, 1500.0, -99.0, -99.0,
        1e-4, 1e-2, "ssrg-rt3_0.19_0.23_1.67", "heymsfield10_particles", -99.0)


def _pam(rows):
    pam = pamtra_amd.pyPamtra()
    for r in rows:
        pam.df.addHydrometeor(r)
    pam = importer.createUsStandardProfile(pam, hgt_lev=np.array([0.0, 1000.0, 2000.0, 3000.0]))
    return pam


def _fullspec_from_dsd(pam):
    """The full spectrum holding the size distribution's own bins."""
    pam.df.addFullSpectra()
    fs = pam.df.dataFullSpec
    for h, row in enumerate(pam.df.data):
        desc = {k: row[k] for k in pam.df.data.dtype.names}
        st, (d, dd, n, rho, deff) = oracle.dsd(desc, pam.nmlSet, np.nan, np.nan, np.nan, 280.0)
        assert st == 0
        bound = np.concatenate([[d[0] - dd[0] / 2], d[0] - dd[0] / 2 + np.cumsum(dd)])
        fs["d_ds"][..., h, :] = deff
        fs["d_bound_ds"][..., h, :] = bound
        fs["n_ds"][..., h, :] = n
        fs["rho_ds"][..., h, :] = rho
        fs["mass_ds"][..., h, :] = desc["a_ms"] * d ** desc["b_ms"] if desc["liq_ice"] == -1 \
            else 1000.0 * np.pi / 6 * d ** 3
        fs["as_ratio"][..., h, :] = desc["as_ratio"]
        fs["canting"][..., h, :] = 0.0
        _, rg = _capi.scattering_model(str(desc["scat_name"]))
        for key, v in zip(("rg_kappa_ds", "rg_beta_ds", "rg_gamma_ds", "rg_zeta_ds"), rg):
            fs[key][..., h, :] = v
    return pam


def _cpu(pam, fullspec):
    nml = dict(passive_nml(pam), hydro_fullspec=fullspec)
    out, st, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, FREQS, nthreads=4,
                            diagnostics=True,
                            descriptorFileFS=pam.df.dataFullSpec if fullspec else None)
    assert st == 0, out["msg"]
    return out


@pytest.mark.parametrize("rows", [[RAIN], [RAIN, SNOW]], ids=["rain", "rain+aggregates"])
def test_fullspec_of_the_dsd_bins_is_the_ordinary_run(rows):
    pam = _pam(rows)
    ref = _cpu(pam, False)
    pam = _fullspec_from_dsd(pam)
    got = _cpu(pam, True)
    np.testing.assert_array_equal(got["nlegen"], ref["nlegen"])
    np.testing.assert_allclose(got["kexttot"], ref["kexttot"], rtol=1e-10)
    np.testing.assert_allclose(got["tb"], ref["tb"], rtol=0, atol=1e-9)


def test_fullspec_empty_bins_leave_the_layer_clear():
    """Bins without particles: the layer has no hydrometeors (threshold > 0)."""
    pam = _fullspec_from_dsd(_pam([RAIN]))
    pam.df.dataFullSpec["n_ds"][0, 0, 1] = 0.0
    out = _cpu(pam, True)
    assert out["nlegen"][0, 0, 1, 0] == -1 and out["nlegen"][0, 0, 0, 0] > 0


def test_fullspec_needs_the_arrays():
    pam = _pam([RAIN])
    with pytest.raises(ValueError):
        _capi.marshal(dict(pam.nmlSet, hydro_fullspec=True), pam.set, pam.df.data,
                      pam.df.data4D, pam.p, {})


def _gpu(pam, fullspec, precision="fp64"):
    from pamtra_amd import _lib
    lib, ctx = _lib.library(), _lib.get_context(0)
    nml = dict(passive_nml(pam), hydro_fullspec=fullspec)
    m = _capi.marshal(nml, dict(pam.set, precision=precision), pam.df.data, pam.df.data4D, pam.p,
                      pam.df.dataFullSpec if fullspec else None)
    f, fp = _capi.freqs_array(FREQS)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    print(f"pam_run fullspec={fullspec} {precision}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    out = dict(hr.arrays)
    out["msg"] = _capi.results_message(hr)
    assert st == 0, out["msg"]
    return out


@pytest.mark.gpu
def test_fullspec_on_the_device_is_the_ordinary_run():
    pam = _pam([RAIN, SNOW])
    ref = _gpu(pam, False)
    pam = _fullspec_from_dsd(pam)
    got = _gpu(pam, True)
    np.testing.assert_array_equal(got["nlegen"], ref["nlegen"])
    np.testing.assert_allclose(got["tb"], ref["tb"], rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_fullspec_with_varying_per_bin_parameters_vs_oracle():
    """Per-bin kappa, beta, gamma, zeta, as_ratio and canting, and a layer
    without particles: the device against the oracle (1e-3 K)."""
    pam = _fullspec_from_dsd(_pam([RAIN, SNOW]))
    fs = pam.df.dataFullSpec
    nb = fs["d_ds"].shape[-1]
    ramp = np.linspace(0.0, 1.0, nb)
    fs["rg_kappa_ds"][..., 1, :] = 0.1 + 0.2 * ramp
    fs["rg_beta_ds"][..., 1, :] = 0.2 + 0.8 * ramp[::-1]
    fs["rg_gamma_ds"][..., 1, :] = 1.5 + 0.3 * ramp
    fs["rg_zeta_ds"][..., 1, :] = 0.5 + 0.5 * ramp
    fs["as_ratio"][..., 1, :] = 0.5 + 0.4 * ramp
    fs["canting"][..., 1, nb // 2:] = 90.0
    fs["n_ds"][0, 0, 2] = 0.0
    g = _gpu(pam, True)
    c = _cpu(pam, True)
    d = np.abs(g["tb"] - c["tb"]).max()
    print(f"max |dTB| = {d:.3e} K", flush=True)
    assert d <= 1e-3
    np.testing.assert_array_equal(g["nlegen"], c["nlegen"])
    mx = _gpu(pam, True, "mixed")
    assert np.abs(mx["tb"] - g["tb"]).max() <= 0.05


@pytest.mark.gpu
def test_fullspec_tmatrix_with_per_bin_axis_ratio_and_canting_vs_oracle():
    """T-matrix hydrometeors from full spectra: per-bin as_ratio and canting
    (scatProperties.f90:261-270), rain oblate, snow with canting in half its
    bins; the device against the oracle (1e-4 K)."""
    pam = _fullspec_from_dsd(_pam([RAIN, SNOW]))
    pam.df.data["scat_name"][:] = "tmatrix"
    fs = pam.df.dataFullSpec
    nb = fs["d_ds"].shape[-1]
    ramp = np.linspace(0.0, 1.0, nb)
    fs["as_ratio"][..., 0, :] = 1.0 - 0.4 * ramp
    fs["as_ratio"][..., 1, :] = 0.5 + 0.4 * ramp
    fs["canting"][..., 1, nb // 2:] = 30.0
    g = _gpu(pam, True)
    c = _cpu(pam, True)
    d = np.abs(g["tb"] - c["tb"]).max()
    print(f"tmatrix fullspec: max |dTB| = {d:.3e} K", flush=True)
    assert d <= 1e-4
    np.testing.assert_array_equal(g["ndoubles"], c["ndoubles"])

