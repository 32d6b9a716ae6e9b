"""The radar simulator's "simple" mode on the device (active runs, the
pyPamtra default): the reference's golden reflectivity and attenuation, and
the oracle on the batched workload for every attenuation mode."""

import ctypes as C
import os

import numpy as np
import pytest

import oracle
from pamtra_amd import _capi, workloads
from _scenario import FREQUENCIES, build_pamtra

pytestmark = pytest.mark.gpu

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "rain_snow_reference.npz"))


def _gpu(pam, freqs, nml, precision="fp64"):
    from pamtra_amd import _lib
    lib = _lib.library()
    ctx = _lib.get_context(0)
    sets = dict(pam.set)
    sets["precision"] = precision
    m = _capi.marshal(nml, sets, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f))
    print(f"pam_run active ncol={m.nx * m.ny} nf={len(f)} {nml.get('radar_attenuation')} "
          f"{nml.get('radar_polarisation')} {precision}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    return dict(hr.arrays), st


def test_golden_radar_through_pypamtra():
    """reference tests/test_regression.py:36-49 on the device: runPamtra with
    the default active=True."""
    pam = build_pamtra()
    assert pam.nmlSet["active"]
    pam.runPamtra(FREQUENCIES)
    np.testing.assert_allclose(pam.r["Ze"], GOLDEN["Ze"], rtol=1e-5)
    np.testing.assert_allclose(pam.r["Att_hydro"], GOLDEN["Att_hydro"], rtol=1e-5)
    np.testing.assert_allclose(pam.r["tb"], GOLDEN["tb"], rtol=1e-5)
    assert (pam.r["Ze"][0, 0, 1:3] > -9999).all()


@pytest.mark.parametrize("att", ["disabled", "top-down", "bottom-up"])
def test_radar_matches_oracle(att):
    pam, freqs = workloads.config2(ncol=12, seed=71)
    nml = dict(pam.nmlSet)
    nml.update(active=True, radar_attenuation=att, radar_polarisation="NN,VV,HV")
    g, st = _gpu(pam, freqs, nml)
    c, st2, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, freqs, nthreads=8)
    assert st == st2 == 0
    assert np.array_equal(g["ze"] == -9999.0, c["ze"] == -9999.0)
    ok = c["ze"] > -9999.0
    assert ok.any()
    d = np.max(np.abs(g["ze"][ok] - c["ze"][ok]))
    print(f"{att}: max |dZe| = {d:.2e} dB over {ok.sum()} values", flush=True)
    assert d <= 1e-6
    np.testing.assert_allclose(g["att_hydro"], c["att_hydro"], rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(g["att_atmo"], c["att_atmo"], rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(g["radar_hgt"], c["radar_hgt"])
    np.testing.assert_array_equal(g["ze"][..., 2, :], -9999.0)  # HV of spheres


def test_radar_mixed_precision_vs_fp64_radar():
    """The mixed path's radar outputs come from its FP32 Mie series (the
    backscattering and extinction sums accumulated in FP64): Ze within 1e-4 dB
    and Att_hydro within 1e-5 relative of the FP64 path; the gas attenuation
    and the radar heights do not involve the Mie stage and are identical."""
    pam, freqs = workloads.config2(ncol=8, seed=72)
    nml = dict(pam.nmlSet)
    nml.update(active=True)
    a, st = _gpu(pam, freqs, nml, "fp64")
    b, st2 = _gpu(pam, freqs, nml, "mixed")
    assert st == st2 == 0
    ok = a["ze"] > -9999.0
    np.testing.assert_array_equal(ok, b["ze"] > -9999.0)
    d = np.abs(a["ze"][ok] - b["ze"][ok]).max()
    print(f"mixed vs fp64: max |dZe| = {d:.2e} dB", flush=True)
    assert d <= 1e-4
    np.testing.assert_allclose(b["att_hydro"], a["att_hydro"], rtol=1e-5, atol=1e-14)
    for k in ("att_atmo", "radar_hgt"):
        np.testing.assert_array_equal(a[k], b[k])


def test_results_netcdf_after_runpamtra(tmp_path):
    """runPamtra's default active+passive results through writeResultsToNetCDF
    (reference core.py:2163-2556), read back with scipy's netCDF-3 reader."""
    from scipy.io import netcdf_file
    pam = build_pamtra()
    pam.runPamtra(FREQUENCIES)
    fn = str(tmp_path / "golden.nc")
    pam.writeResultsToNetCDF(fn)
    with netcdf_file(fn, "r", mmap=False) as nc:
        np.testing.assert_array_equal(nc.variables["tb"].data, pam.r["tb"].astype(np.float32))
        np.testing.assert_array_equal(nc.variables["Ze"].data, pam.r["Ze"].astype(np.float32))
        np.testing.assert_array_equal(nc.variables["height"].data,
                                      pam.r["radar_hgt"].astype(np.float32))
        np.testing.assert_allclose(nc.variables["tb"].data, GOLDEN["tb"], rtol=1e-5)


def test_run_parallel_pamtra_is_runpamtra():
    """runParallelPamtra (reference core.py:1570-1675) gives runPamtra's results."""
    pam = build_pamtra()
    pam.runParallelPamtra(FREQUENCIES)
    np.testing.assert_allclose(pam.r["tb"], GOLDEN["tb"], rtol=1e-5)
    np.testing.assert_allclose(pam.r["Ze"], GOLDEN["Ze"], rtol=1e-5)
