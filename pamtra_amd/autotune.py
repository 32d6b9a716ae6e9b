"""Autotuner of the launch shapes of a context (pam_context_set_tuning).

The reference has no launch shape to choose (its Fortran loops run one
column at a time, python/pyPamtra/core.py:1518-1568).  Here the three heavy
kernels can run one workgroup per unit or as a persistent grid of k
workgroups per compute unit taking units from a work queue, and a
single-chunk call can be captured in a HIP graph and replayed.  Which is
faster depends on the device and on the workload (units per call, their
spread of cost), so the choice is measured:

* persistent grids: every candidate value is applied to the three kernels
  at once and each kernel keeps the value that gave its own stage the
  shortest time (the stage times come from HIP events, pam_last_stage_times);
* graph replay and pipelining (the call's column chunks on two streams,
  each chunk's DSD and Mie beside the previous chunk's radiative transfer):
  timed on the whole call, each against the plain single-stream launch.

Every choice gives bitwise the same results; only time changes.  Choices are
cached per device name and workload signature in a JSON file.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import time
from typing import Callable, Dict, Optional, Sequence

PERSISTENT_KEYS = ("mie_persistent", "doubling_persistent", "adding_persistent")
STAGE_OF_KEY = {"mie_persistent": "mie", "doubling_persistent": "doubling",
                "adding_persistent": "adding"}
PERSISTENT_CANDIDATES = (0, 1, 2, 4)
PIPELINE = 4  # column chunks of a pipelined call
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pamtra_amd", "tuning.json")


def get_tuning(lib, ctx) -> Dict[str, int]:
    keys = lib.pam_tuning_keys().decode().split(",")
    return {k: int(lib.pam_context_get_tuning(ctx, k.encode())) for k in keys}


def apply(lib, ctx, choice: Dict[str, int]) -> None:
    for k, v in choice.items():
        if k.startswith("_"):
            continue
        if lib.pam_context_set_tuning(ctx, k.encode(), int(v)) != 0:
            raise ValueError(f"pam_context_set_tuning rejected {k}={v}")


def _stage_times(lib, ctx) -> Dict[str, float]:
    names = lib.pam_stage_names().decode().split(",")
    ms = (C.c_double * len(names))()
    lib.pam_last_stage_times(ctx, ms, len(names))
    return dict(zip(names, ms[:]))


def tune(lib, ctx, run_once: Callable[[], None], repeats: int = 2,
         candidates: Sequence[int] = PERSISTENT_CANDIDATES,
         timer: Callable[[], float] = time.perf_counter,
         stage_times: Optional[Callable[[], Dict[str, float]]] = None) -> Dict[str, object]:
    """Measures the candidates with `run_once` (one synchronous call of the
    workload on `ctx`), applies the fastest and returns it (with the
    measured times under "_ms")."""
    stage_times = stage_times or (lambda: _stage_times(lib, ctx))
    # the stage times of a pipelined call overlap: shapes are chosen on one stream
    apply(lib, ctx, {"graph": 0, "pipeline": 0})
    run_once()  # warm: workspace, tables
    best: Dict[str, tuple] = {}
    table: Dict[str, Dict[int, float]] = {k: {} for k in PERSISTENT_KEYS}
    for v in candidates:
        apply(lib, ctx, {k: v for k in PERSISTENT_KEYS})
        acc = {k: 0.0 for k in PERSISTENT_KEYS}
        for _ in range(repeats):
            run_once()
            st = stage_times()
            for k in PERSISTENT_KEYS:
                acc[k] += st[STAGE_OF_KEY[k]]
        for k in PERSISTENT_KEYS:
            table[k][v] = acc[k] / repeats
            if k not in best or table[k][v] < best[k][1]:
                best[k] = (v, table[k][v])
    choice = {k: best[k][0] for k in PERSISTENT_KEYS}
    apply(lib, ctx, choice)
    whole = {}
    for g, pl in ((0, 0), (1, 0), (0, PIPELINE)):
        apply(lib, ctx, {"graph": g, "pipeline": pl})
        run_once()  # capture (graph) or warm
        t0 = timer()
        for _ in range(repeats):
            run_once()
        whole[(g, pl)] = (timer() - t0) / repeats * 1e3
    choice["graph"], choice["pipeline"] = min(whole, key=whole.get)
    apply(lib, ctx, {"graph": choice["graph"], "pipeline": choice["pipeline"]})
    out: Dict[str, object] = dict(choice)
    out["_ms"] = {"persistent": {k: {str(v): round(t, 3) for v, t in table[k].items()}
                                 for k in PERSISTENT_KEYS},
                  "whole": {f"graph={g},pipeline={pl}": round(t, 3)
                            for (g, pl), t in whole.items()}}
    return out


def device_name() -> str:
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # no device query available: one shared entry
        return "unknown"


def cached_tune(lib, ctx, run_once: Callable[[], None], signature: str,
                cache_path: Optional[str] = DEFAULT_CACHE, retune: bool = False,
                **kw) -> Dict[str, object]:
    """`tune` with the result cached under (device name, signature)."""
    key = f"{device_name()}|{signature}"
    cache: Dict[str, Dict[str, object]] = {}
    if cache_path and os.path.exists(cache_path):
        try:
            with open(cache_path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if not retune and key in cache:
        choice = {k: int(v) for k, v in cache[key].items() if not k.startswith("_")}
        apply(lib, ctx, choice)
        return dict(cache[key], _cached=True)
    res = tune(lib, ctx, run_once, **kw)
    if cache_path:
        try:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
            cache[key] = res
            tmp = cache_path + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(cache, fh, indent=1, sort_keys=True)
            os.replace(tmp, cache_path)
        except OSError:
            pass
    return res
