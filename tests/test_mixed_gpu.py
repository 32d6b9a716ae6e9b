"""The mixed-precision path (set["precision"] = "mixed") against the FP64
path on the same inputs: per channel, max |dTB| <= 0.05 K."""

import ctypes as C

import numpy as np
import pytest

from pamtra_amd import _capi, workloads

pytestmark = pytest.mark.gpu

MIXED_TOL = 0.05  # K, mixed precision vs FP64


def _run(pam, freqs, precision):
    from pamtra_amd import _lib
    lib = _lib.library()
    ctx = _lib.get_context(0)
    sets = dict(pam.set)
    sets["precision"] = precision
    nml = dict(pam.nmlSet)
    nml["active"] = False
    m = _capi.marshal(nml, sets, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    print(f"pam_run {precision} ncol={m.nx * m.ny} nf={len(f)}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    return dict(hr.arrays), st


def _per_channel(a, b, freqs):
    d = np.abs(a["tb"] - b["tb"])  # (x, y, lev, angle, freq, stokes)
    per = [float(np.max(d[..., k, :])) for k in range(len(freqs))]
    print("max |dTB| per channel (K):",
          ", ".join(f"{f:g} GHz {v:.2e}" for f, v in zip(freqs, per)), flush=True)
    return per


@pytest.mark.parametrize("cfg,ncol", [("config2", 48), ("config3", 32)])
def test_mixed_matches_fp64_per_channel(cfg, ncol):
    pam, freqs = getattr(workloads, cfg)(ncol=ncol)
    a, st = _run(pam, freqs, "fp64")
    b, st2 = _run(pam, freqs, "mixed")
    assert st == 0 and st2 == 0
    # the same layers scatter; the mixed path's higher-order start layer
    # needs fewer doublings than the reference's first-order one
    na, nb = a["ndoubles"], b["ndoubles"]
    np.testing.assert_array_equal(na < 0, nb < 0)
    assert np.all(nb[na >= 0] <= na[na >= 0])
    assert np.mean(nb[na >= 0]) < np.mean(na[na >= 0]) - 8
    per = _per_channel(a, b, freqs)
    assert max(per) <= MIXED_TOL
    # the FP32 Mie series: extinction per (layer, channel) against the FP64
    # series, within 1e-5 relative, plus what the skipped bins may carry: the
    # mixed path of a passive run skips a bin whose extinction is bounded
    # below 1e-13 m^-1, so a layer of four hydrometeors of 100 bins loses at
    # most 4e-11 m^-1
    ka, kb = a["kexttot"], b["kexttot"]
    sc = ka > 0
    np.testing.assert_array_equal(sc, kb > 0)
    nh = len(pam.df.data)
    slack = nh * int(np.max(pam.df.data["nbin"])) * 1e-13
    rk = np.abs(kb[sc] / ka[sc] - 1.0).max()
    ex = (np.abs(kb[sc] - ka[sc]) - 1e-5 * ka[sc] - slack).max()
    print(f"max relative dkext = {rk:.2e}; excess over 1e-5 rel + {slack:.1e} = {ex:.2e}",
          flush=True)
    assert ex <= 0.0


def test_mixed_serving_mix_high_frequency():
    """325 GHz and 13 mm particles: the thickest, most scattering layers."""
    (pam, freqs), = workloads.serving_requests(1, seed=21, max_cols=24)
    freqs = sorted(set(freqs) | {183.31, 325.0})
    a, st = _run(pam, freqs, "fp64")
    b, st2 = _run(pam, freqs, "mixed")
    assert st == 0 and st2 == 0
    assert max(_per_channel(a, b, freqs)) <= MIXED_TOL


def _bins(lib, ctx):
    work = (C.c_int64 * 14)()
    lib.pam_last_work(ctx, work, 14)
    return int(work[6])


def test_mixed_skips_negligible_bins_in_passive_runs():
    """The four-category columns carry exponential tails up to 13 mm whose
    largest bins hold almost no particles: the mixed path of a passive run
    computes fewer Mie bins than the FP64 path (which computes every bin with
    particles, as the reference does), and an active run computes them all."""
    from pamtra_amd import _lib
    lib, ctx = _lib.library(), _lib.get_context(0)
    pam, freqs = workloads.config3(ncol=16)
    _run(pam, freqs, "fp64")
    n64 = _bins(lib, ctx)
    _run(pam, freqs, "mixed")
    nmix = _bins(lib, ctx)
    print(f"Mie bins: fp64 {n64}, mixed {nmix}", flush=True)
    assert 0 < nmix < n64
    sets = dict(pam.set, precision="mixed")
    nml = dict(pam.nmlSet, active=True)
    m = _capi.marshal(nml, sets, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f))
    assert lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                       C.byref(m.profile), fp, len(f), C.byref(hr.struct)) == 0
    assert _bins(lib, ctx) == n64
