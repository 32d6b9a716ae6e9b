"""The request loop on the device: each request's result is the same alone
and in the mix (the serving contract), and matches the oracle."""

import numpy as np
import pytest

import oracle
from pamtra_amd import serving, workloads

pytestmark = pytest.mark.gpu


def _reqs(n, max_cols, seed):
    out = []
    for i, (pam, freqs) in enumerate(workloads.serving_requests(n, seed=seed, max_cols=max_cols)):
        out.append(serving.Request(pam=pam, freqs=freqs, arrival=0.002 * i, rid=i))
    return out


def test_alone_equals_in_the_mix():
    reqs = _reqs(8, 48, seed=2024)
    print("requests (columns, layers, channels):",
          [(r.columns, int(r.pam.p["max_nlyrs"]), len(r.freqs)) for r in reqs], flush=True)
    alone = []
    with serving.ForwardServer(slots=1, hbm_budget_gb=8.0) as srv:
        for r in reqs:
            (res,) = srv.run([r], realtime=False)
            assert res.status == 0, res.message
            alone.append(res.tb)
    with serving.ForwardServer(slots=4, hbm_budget_gb=8.0) as srv:
        mixed = srv.run(reqs, realtime=True)
    assert max(r.slot for r in mixed) > 0  # the mix did use several slots
    for a, m in zip(alone, mixed):
        assert m.status == 0, m.message
        np.testing.assert_array_equal(a, m.tb)


def test_served_request_matches_the_oracle():
    (req,) = _reqs(1, 6, seed=77)
    with serving.ForwardServer(slots=2, hbm_budget_gb=4.0) as srv:
        (res,) = srv.run([req], realtime=False)
    assert res.status == 0, res.message
    pam = req.pam
    nml = dict(pam.nmlSet)
    ref, st, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, req.freqs,
                            nthreads=8)
    assert st == 0
    d = np.max(np.abs(res.tb - ref["tb"]))
    print(f"served vs oracle max |dTB| = {d:.3e} K", flush=True)
    assert d <= 1e-3


def test_tight_budget_serialises_but_completes():
    reqs = _reqs(6, 32, seed=5)
    with serving.ForwardServer(slots=4, hbm_budget_gb=1.0) as srv:
        need = max(srv.need_bytes(r) for r in reqs)
        srv.budget = int(need * 9 // 8 * 1.5)  # room for one request at a time
        res = srv.run(reqs, realtime=False)
        assert srv.peak_reserved <= srv.budget
    assert all(r.status == 0 for r in res)


def test_batching_server_packs_requests_and_keeps_each_result():
    """Continuous batching: the requests waiting together share one call per
    channel (fewer calls than request-channels), and every request's TB is
    bitwise its result alone."""
    reqs = _reqs(10, 40, seed=2025)
    alone = []
    with serving.ForwardServer(slots=1, hbm_budget_gb=8.0) as srv:
        for r in reqs:
            (res,) = srv.run([r], realtime=False)
            assert res.status == 0, res.message
            alone.append(res.tb)
    with serving.BatchingServer(workers=2, hbm_budget_gb=8.0) as bs:
        batched = bs.run(reqs, realtime=False)
        calls = bs.calls
    print(f"{calls} calls for {sum(len(r.freqs) for r in reqs)} request-channels", flush=True)
    assert calls < sum(len(r.freqs) for r in reqs)
    for a, m in zip(alone, batched):
        assert m.status == 0, m.message
        np.testing.assert_array_equal(a, m.tb)
