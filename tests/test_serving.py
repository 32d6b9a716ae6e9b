"""The request loop's admission and bookkeeping (pamtra_amd.serving), on the
CPU with a stand-in for the device call."""

import threading
import time
import types

import numpy as np
import pytest

from pamtra_amd import serving


class _Fake:
    """A device call that sleeps, records concurrency and the budget in use."""

    def __init__(self, server_ref, dur=0.02):
        self.server_ref = server_ref
        self.dur = dur
        self.lock = threading.Lock()
        self.running = 0
        self.max_running = 0
        self.max_held = 0

    def __call__(self, slot, req):
        with self.lock:
            self.running += 1
            self.max_running = max(self.max_running, self.running)
            self.max_held = max(self.max_held, sum(self.server_ref[0]._held))
        time.sleep(self.dur * (1 + req.rid % 3))
        with self.lock:
            self.running -= 1
        return 0, "", np.full(3, float(req.rid))


def _requests(n, rate=None, seed=0):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate else np.zeros(n)
    pam = types.SimpleNamespace(p={"ngridx": 2, "ngridy": 1})
    return [serving.Request(pam=pam, freqs=[1.0], arrival=float(t[i]), rid=i)
            for i in range(n)]


def _server(slots, budget_gb, need, dur=0.02):
    ref = [None]
    fake = _Fake(ref, dur)
    srv = serving.ForwardServer(slots=slots, hbm_budget_gb=budget_gb, runner=fake, need=need)
    ref[0] = srv
    return srv, fake


def test_every_request_completes_with_its_own_result_and_order():
    srv, fake = _server(4, 1.0, need=lambda r: 1 << 20)
    reqs = _requests(20)
    res = srv.run(reqs, realtime=False)
    assert [r.rid for r in res] == list(range(20))
    for r in res:
        assert r.status == 0 and r.tb[0] == r.rid
    assert fake.max_running == 4


def test_admission_never_exceeds_the_hbm_budget():
    """Requests of 0.3 GiB under a 1 GiB budget: at most 2 run at once
    (3 x 0.3 x 9/8 > 1), although 4 slots are free."""
    gib = 1 << 30
    srv, fake = _server(4, 1.0, need=lambda r: int(0.3 * gib))
    res = srv.run(_requests(12), realtime=False)
    assert all(r.status == 0 for r in res)
    assert fake.max_running == 2
    assert fake.max_held <= gib and srv.peak_reserved <= gib


def test_a_request_over_the_budget_is_refused_before_any_work():
    srv, fake = _server(2, 0.001, need=lambda r: 1 << 30)
    with pytest.raises(ValueError):
        srv.run(_requests(2), realtime=False)
    assert fake.max_running == 0


def test_latency_is_counted_from_arrival_and_summarised():
    srv, _ = _server(2, 1.0, need=lambda r: 1 << 20, dur=0.01)
    reqs = _requests(30, rate=200.0, seed=3)
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    for r, q in zip(res, reqs):
        assert r.start >= q.arrival - 1e-3 and r.end >= r.start
        assert r.latency >= 0.01
    s = serving.summarize(res, reqs, wall)
    assert s["requests"] == 30 and s["failed"] == 0 and s["columns_channels"] == 60
    assert 0 < s["p50_ms"] <= s["p99_ms"]


def test_a_failing_call_is_reported_and_the_loop_goes_on():
    def runner(slot, req):
        if req.rid == 1:
            raise RuntimeError("boom")
        return 0, "", np.zeros(1)
    srv = serving.ForwardServer(slots=2, hbm_budget_gb=1.0, runner=runner,
                                need=lambda r: 1 << 20)
    res = srv.run(_requests(4), realtime=False)
    assert res[1].status == 2 and "boom" in res[1].message
    assert all(res[i].status == 0 for i in (0, 2, 3))


def test_merged_columns_give_each_request_its_own_result():
    """merge_columns: requests of different layer counts concatenated along x
    and padded give, through the oracle, exactly each request's own TB."""
    import oracle
    from pamtra_amd import workloads
    mix = workloads.serving_requests(3, seed=31, max_cols=3)
    pams = [p for p, _ in mix]
    assert len({int(p.p["max_nlyrs"]) for p in pams}) > 1
    ch = [36.5]
    prof, d4 = serving.merge_columns(pams)
    assert prof["ngridx"] == sum(int(p.p["ngridx"]) for p in pams)
    nml = dict(pams[0].nmlSet)
    nml["active"] = False
    merged, st, _ = oracle.run(nml, pams[0].set, pams[0].df.data, d4, prof, ch, nthreads=4)
    assert st == 0
    o = 0
    for p in pams:
        alone, st2, _ = oracle.run(nml, p.set, p.df.data, p.df.data4D, p.p, ch, nthreads=4)
        assert st2 == 0
        n = int(p.p["ngridx"])
        np.testing.assert_array_equal(merged["tb"][o:o + n], alone["tb"])
        o += n


def test_group_key_separates_settings_and_descriptors():
    from pamtra_amd import workloads
    (a, _), (b, _) = workloads.serving_requests(2, seed=32, max_cols=2)
    assert serving._group_key(a) == serving._group_key(b)
    b.set["precision"] = "mixed"
    assert serving._group_key(a) != serving._group_key(b)
