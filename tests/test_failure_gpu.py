"""Failure detection on the device: each failure the reference stops on is
confined to its (column, channel), reported with the reference's message,
and every other (column, channel) keeps its result bit for bit.

* CHECK_NORM4 (radscat4.f90:125-131): reached naturally by dense graupel and
  snow at 183-325 GHz, where the 16-angle quadrature no longer normalises the
  phase matrix; the device flags the same units as the oracle.
* NaN Legendre coefficients (scatcnv.f90:38-43) and a singular (I - R R)
  (radmat.f90:180-189, where the reference STOPs the process): forced on one
  named (column, channel) from the host (PAM_TEST_FAULT), every kernel still
  exiting normally."""

import ctypes as C

import numpy as np
import pytest

import oracle
from pamtra_amd import _capi, workloads

pytestmark = pytest.mark.gpu

FREQS = [89.0, 183.31, 325.0]


def _gpu(pam, freqs):
    from pamtra_amd import _lib
    lib, ctx = _lib.library(), _lib.get_context(0)
    nml = dict(pam.nmlSet)
    nml["active"] = False
    m = _capi.marshal(nml, pam.set, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f))
    print(f"pam_run ncol={m.nx * m.ny} nf={len(f)}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    return st, _capi.results_message(hr), hr.arrays["tb"], hr.arrays["status"]


def _dense(ncol, seed, q, cols):
    pam, _ = workloads.config3(ncol=ncol, seed=seed)
    for h in (2, 3):  # snow, graupel
        sel = np.zeros(pam.p["hydro_q"].shape[:3], dtype=bool)
        sel[list(cols)] = True
        a = pam.p["hydro_q"][..., h]
        pam.p["hydro_q"][..., h] = np.where(sel & (a > 0), q, a)
    return pam


def test_norm_violation_matches_the_oracle_and_stays_in_its_units():
    pam = _dense(6, 5, 5e-3, cols=[1, 3, 5])
    st, msg, tb, status = _gpu(pam, FREQS)
    nml = dict(pam.nmlSet)
    nml["active"] = False
    ref, st2, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, FREQS, nthreads=8)
    assert st == st2 == _capi.PAM_FATAL
    assert "not normalized" in msg and "not normalized" in ref["msg"], (msg, ref["msg"])
    np.testing.assert_array_equal(status, ref["status"])
    bad = status == _capi.PAM_FATAL
    assert bad[[3, 5], 0, 2].all() and not bad[[0, 1, 2, 4]].any()
    ok = np.isfinite(ref["tb"])
    np.testing.assert_array_equal(ok, np.isfinite(tb))
    assert np.max(np.abs(tb[ok] - ref["tb"][ok])) <= 1e-3
    # the columns without a violation, run without the others, give the
    # same TB bit for bit
    keep = [0, 1, 2, 4]
    clean = _dense(6, 5, 5e-3, cols=[1, 3, 5])
    for k, v in list(clean.p.items()):
        if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] == 6:
            clean.p[k] = v[keep]
    for k, v in list(clean.df.data4D.items()):
        clean.df.data4D[k] = v[keep]
    clean.p["ngridx"] = len(keep)
    st3, _, tb3, status3 = _gpu(clean, FREQS)
    assert st3 == 0 and not status3.any()
    np.testing.assert_array_equal(tb3, tb[keep])


@pytest.mark.parametrize("kind,col,ch,text", [("nan_coef", 3, 2, "nan in Legendre"),
                                              ("singular", 5, 1, "Matrix inversion failed")])
def test_forced_failure_is_confined_to_its_unit(monkeypatch, kind, col, ch, text):
    pam, freqs = workloads.config2(ncol=8, seed=41)
    freqs = freqs[:4]
    st0, _, tb0, status0 = _gpu(pam, freqs)
    assert st0 in (0, 1)
    monkeypatch.setenv("PAM_TEST_FAULT", f"{kind}:{col}:{ch}")
    print(f"PAM_TEST_FAULT={kind}:{col}:{ch}", flush=True)
    st, msg, tb, status = _gpu(pam, freqs)
    monkeypatch.delenv("PAM_TEST_FAULT")
    assert st == _capi.PAM_FATAL and text in msg, msg
    assert status[col, 0, ch] == _capi.PAM_FATAL
    assert np.all(np.isnan(tb[col, 0, ..., ch, :]))
    other = np.ones(status.shape, dtype=bool)
    other[col, 0, ch] = False
    np.testing.assert_array_equal(status[other], status0[other])
    mask = np.ones(tb.shape, dtype=bool)
    mask[col, 0, ..., ch, :] = False
    np.testing.assert_array_equal(tb[mask], tb0[mask])
