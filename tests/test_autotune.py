"""Selection logic of pamtra_amd.autotune against a stand-in library and
timer (no device)."""

import json

from pamtra_amd import autotune


class FakeLib:
    """pam_tuning_* and pam_last_stage_times of a context whose stage times
    follow a given cost model of the tuning values."""

    KEYS = b"mie_persistent,doubling_persistent,adding_persistent,graph,pipeline"

    def __init__(self, cost):
        self.tune = {k: 0 for k in self.KEYS.decode().split(",")}
        self.cost = cost
        self.calls = 0

    def pam_tuning_keys(self):
        return self.KEYS

    def pam_context_set_tuning(self, ctx, key, value):
        k = key.decode()
        if k not in self.tune or value < 0 or value > 64 or (k == "graph" and value > 1) or (
                k == "pipeline" and value > 16):
            return 2
        self.tune[k] = value
        return 0

    def pam_context_get_tuning(self, ctx, key):
        return self.tune.get(key.decode(), -1)

    def stage_times(self):
        return {st: self.cost(st, self.tune) for st in ("mie", "doubling", "adding")}


def _setup(cost):
    lib = FakeLib(cost)
    clock = [0.0]

    def run_once():
        lib.calls += 1
        # graph replay saves launch time; pipelining overlaps 30 % of the stages
        t = sum(lib.stage_times().values()) * 1e-3 * (0.7 if lib.tune["pipeline"] else 1.0)
        clock[0] += t + (0.0 if lib.tune["graph"] else 0.002)

    return lib, clock, run_once


def test_picks_fastest_per_stage_and_graph():
    best = {"mie": 2, "doubling": 4, "adding": 0}
    key = {"mie": "mie_persistent", "doubling": "doubling_persistent",
           "adding": "adding_persistent"}

    def cost(st, tune):
        return 10.0 + abs(tune[key[st]] - best[st])

    lib, clock, run_once = _setup(cost)
    res = autotune.tune(lib, None, run_once, repeats=1, timer=lambda: clock[0],
                        stage_times=lib.stage_times)
    assert res["mie_persistent"] == 2
    assert res["doubling_persistent"] == 4
    assert res["adding_persistent"] == 0
    assert res["graph"] == 0 and res["pipeline"] == autotune.PIPELINE
    # the choice is left applied on the context
    assert autotune.get_tuning(lib, None) == {k: res[k] for k in lib.tune}


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "device_name", lambda: "dev")
    lib, clock, run_once = _setup(lambda st, tune: 5.0 + tune["mie_persistent"])
    path = str(tmp_path / "t.json")
    r1 = autotune.cached_tune(lib, None, run_once, "sig", cache_path=path, repeats=1,
                              timer=lambda: clock[0], stage_times=lib.stage_times)
    n1 = lib.calls
    assert r1["mie_persistent"] == 0
    assert "dev|sig" in json.load(open(path))
    lib.tune["mie_persistent"] = 4
    r2 = autotune.cached_tune(lib, None, run_once, "sig", cache_path=path)
    assert r2["_cached"] and lib.calls == n1          # no runs: from the cache
    assert lib.tune["mie_persistent"] == 0            # and applied


def test_rejected_value_raises():
    lib, _, _ = _setup(lambda st, tune: 1.0)
    try:
        autotune.apply(lib, None, {"graph": 3})
    except ValueError:
        return
    raise AssertionError("an invalid value must raise")


def test_graph_when_pipelining_does_not_pay():
    lib, clock, run_once = _setup(lambda st, tune: 0.001)
    res = autotune.tune(lib, None, run_once, repeats=1, timer=lambda: clock[0],
                        stage_times=lib.stage_times)
    assert res["graph"] == 1 and res["pipeline"] == 0
