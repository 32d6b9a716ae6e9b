"""The 32-angle library (libpamtra_amd_n64.so: 32 quadrature angles per
hemisphere, 64x64 reflection/transmission operators), as a reference rebuilt
with nummu = 32 (settings.f90:16) would run it, against the 32-angle oracle:
the large-domain configuration's columns, the golden scenario through
runPamtra, interior output levels and the inverse fallbacks.  FP64 bar
|dTB| <= 1e-3 K; mixed <= 0.05 K against FP64."""

import ctypes as C

import numpy as np
import pytest

import oracle
from pamtra_amd import _capi, workloads
from _scenario import FREQUENCIES, build_pamtra, passive_nml

pytestmark = pytest.mark.gpu

CH6 = [10.65, 23.8, 89.0, 166.0, 183.31, 325.0]


def _gpu(pam, freqs, precision="fp64"):
    from pamtra_amd import _lib
    sets = dict(pam.set, _nangles=pam._nangles, precision=precision)
    m = _capi.marshal(passive_nml(pam), sets, pam.df.data, pam.df.data4D, pam.p)
    assert m.nummu == 32
    lib = _lib.library(32)
    ctx = _lib.get_context(0, 32)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    print(f"pam_run 32 angles ncol={m.nx * m.ny} L={m.max_nlyrs} nf={len(f)} {precision}",
          flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    out = dict(hr.arrays)
    out["msg"] = _capi.results_message(hr)
    return out, st


def _cpu(pam, freqs):
    out, st, _ = oracle.run(passive_nml(pam), dict(pam.set, _nangles=32), pam.df.data,
                            pam.df.data4D, pam.p, freqs, nthreads=16, diagnostics=True)
    return out, st


def _compare(g, c, tol=1e-3):
    assert g["tb"].shape[3] == 64
    ok = np.isfinite(c["tb"])
    assert ok.all() and np.array_equal(ok, np.isfinite(g["tb"]))
    d = np.abs(g["tb"] - c["tb"])
    print(f"max |dTB| = {d.max():.3e} K", flush=True)
    assert d.max() <= tol
    np.testing.assert_array_equal(g["ndoubles"], c["ndoubles"])
    np.testing.assert_array_equal(g["nlegen"], c["nlegen"])


def test_config5_columns_vs_oracle():
    """150 layers, four categories of 100 bins, all twelve channels
    (10.65-325 GHz), ocean and land."""
    pam, freqs = workloads.config5(ncol=3, seed=1005)
    assert len(freqs) == 12
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    np.testing.assert_allclose(g["emissivity"], c["emissivity"], rtol=1e-12, atol=1e-15)
    _compare(g, c)


def test_config5_mixed_precision_vs_fp64():
    pam, freqs = workloads.config5(ncol=24, seed=1006)
    a, st = _gpu(pam, freqs)
    b, st2 = _gpu(pam, freqs, "mixed")
    assert st == 0 and st2 == 0
    d = np.abs(a["tb"] - b["tb"]).max(axis=(0, 1, 2, 3, 5))
    print("mixed vs fp64 max |dTB| per channel:", np.array2string(d, precision=5), flush=True)
    assert np.all(d <= 0.05)


def test_golden_scenario_through_pypamtra_at_32_angles():
    pam = build_pamtra()
    pam._nangles = 32
    pam.nmlSet["active"] = False
    pam.runPamtra(FREQUENCIES)
    assert pam.r["tb"].shape == (1, 1, 2, 64, 2, 2)
    assert pam.r["angles_deg"].shape == (64,)
    c, st = _cpu(pam, FREQUENCIES)
    assert st == 0
    np.testing.assert_allclose(pam.r["tb"], c["tb"], rtol=0, atol=1e-3)


@pytest.mark.parametrize("mode", [None, "gauss-jordan"])
def test_interior_levels_and_inverse_modes_at_32_angles(monkeypatch, mode):
    pam, _ = workloads.config5(ncol=2, seed=1007)
    hl = pam.p["hgt_lev"][:, 0, :]
    nx, ny = pam.p["ngridx"], pam.p["ngridy"]
    pam.p["noutlevels"] = 3
    pam.p["obs_height"] = np.array([[833000.0, 0.0, hl[c, 40]] for c in range(2)]).reshape(nx, ny, 3)
    freqs = [31.4, 183.31]
    if mode:
        monkeypatch.setenv("PAM_INVERSE", mode)
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    _compare(g, c)
