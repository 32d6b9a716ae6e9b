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
    # the FP32 Mie series: extinction and single-scattering albedo per
    # (layer, channel) against the FP64 series
    ka, kb = a["kexttot"], b["kexttot"]
    sc = ka > 0
    np.testing.assert_array_equal(sc, kb > 0)
    rk = np.abs(kb[sc] / ka[sc] - 1.0).max()
    print(f"max relative dkext = {rk:.2e}", flush=True)
    assert rk <= 1e-5


def test_mixed_serving_mix_high_frequency():
    """325 GHz and 13 mm particles: the thickest, most scattering layers."""
    (pam, freqs), = workloads.serving_requests(1, seed=21, max_cols=24)
    freqs = sorted(set(freqs) | {183.31, 325.0})
    a, st = _run(pam, freqs, "fp64")
    b, st2 = _run(pam, freqs, "mixed")
    assert st == 0 and st2 == 0
    assert max(_per_channel(a, b, freqs)) <= MIXED_TOL
