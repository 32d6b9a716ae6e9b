"""Request loop of the passive forward operator on one device.

The reference has no serving mode: each ``runPamtra`` call allocates its
module globals, runs every (column, channel) and returns
(``python/pyPamtra/core.py:1518-1568``); its only concurrency is one process
per core (``core.py:1613-1647``).  Here a server keeps ``slots`` contexts of
the C-ABI on one GPU, each with its own non-blocking stream and one host
thread, so requests that arrive while others are on the device run
alongside them:

* admission: a request starts when a slot is free and the device workspace
  of all slots, with this request's need (``pam_workspace_bytes``), fits the
  HBM budget; idle slots give their workspace back (``pam_context_trim``)
  when the budget is short;
* the Mie angle tables are built once per device and shared by every slot;
* each request is timed from its arrival to its result on the host
  (latency), and the server reports throughput and latency percentiles.

A request's result does not depend on what runs beside it: every unit of
work is a (column, channel) of that request, computed by the same kernels on
its own stream and workspace.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
import queue
import threading
import time
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import _capi, _lib


@dataclasses.dataclass
class Request:
    """One forward-operator call: a pyPamtra-like object (``p``, ``nmlSet``,
    ``set``, ``df``) and its channels; ``arrival`` is seconds after the start
    of ``ForwardServer.run``."""
    pam: object
    freqs: Sequence[float]
    arrival: float = 0.0
    rid: int = 0

    @property
    def columns(self) -> int:
        return int(self.pam.p["ngridx"]) * int(self.pam.p["ngridy"])


@dataclasses.dataclass
class Result:
    rid: int
    status: int
    message: str
    tb: np.ndarray
    arrival: float
    start: float
    end: float
    slot: int
    need_bytes: int

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def scattering_layers_bound(pam) -> int:
    """Upper bound of the layers that hold hydrometeors (the device's work
    items): layers inside nlyrs with any positive content."""
    p = pam.p
    nl = np.asarray(p["nlyrs"]).reshape(-1)
    L = int(p["max_nlyrs"])
    inside = np.arange(L)[None, :] < nl[:, None]
    any_h = np.zeros_like(inside)
    for key in ("hydro_q", "hydro_n", "hydro_reff"):
        a = p.get(key)
        if isinstance(a, np.ndarray) and a.size:
            a = a.reshape(nl.size, L, -1)
            any_h |= np.any(a > 0.0, axis=-1)
    return int(np.count_nonzero(any_h & inside))


def workspace_bytes(lib, pam, nfreq: int, optics_limit: Optional[int] = None) -> int:
    """Device bytes of one call of this request (the library's own sizing).
    With `optics_limit` (the context's workspace limit) the layers are
    processed in chunks of at most that many bytes of optics."""
    p = pam.p
    ncol = int(p["ngridx"]) * int(p["ngridy"])
    nbins = [2 if str(r["dist_name"]).strip() == "mono" else int(r["nbin"])
             for r in pam.df.data] or [2]
    nb = max(nbins)
    if "nbin" in pam.df.data4D:
        nb = max(nb, int(np.max(pam.df.data4D["nbin"])))
    dims = (ncol, int(p["max_nlyrs"]), len(pam.df.data), nb, nfreq, int(p["noutlevels"]))
    nitems = scattering_layers_bound(pam)
    if optics_limit is not None:
        per_item = lib.pam_workspace_bytes(*dims, 2) - lib.pam_workspace_bytes(*dims, 1)
        nitems = min(nitems, max(1, optics_limit // max(1, per_item)))
    n = lib.pam_workspace_bytes(*dims, nitems)
    if n < 0:
        raise ValueError("pam_workspace_bytes: invalid dimensions")
    return int(n)


def _run_on_context(lib, ctx, req: Request):
    """One pam_run of the request on a slot's context; returns (status, msg, tb)."""
    pam = req.pam
    nml = dict(pam.nmlSet)
    nml["active"] = False
    m = _capi.marshal(nml, pam.set, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(req.freqs)
    hr = _capi.alloc_results(m, len(f))
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    return int(st), _capi.results_message(hr), hr.arrays["tb"]


class ForwardServer:
    """``slots`` concurrent contexts on ``device`` under an HBM budget.

    ``runner(slot, request) -> (status, message, tb)`` replaces the device
    call (used by the CPU tests of the admission logic); ``need(request)``
    replaces the workspace sizing then.  ``tuning`` (pam_context_set_tuning
    keys, e.g. from pamtra_amd.autotune) is applied to every slot's context.
    """

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget_gb: float = 64.0,
                 runner: Optional[Callable] = None, need: Optional[Callable] = None,
                 tuning: Optional[dict] = None):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.slots = slots
        self.budget = int(hbm_budget_gb * (1 << 30))
        # each slot processes a request's layers in chunks of at most this
        # much optics, so that every request fits a slot's share
        self.optics_limit = int(0.75 * self.budget / slots)
        self._runner = runner
        self._need = need
        self._lib = None
        self._ctx: List[C.c_void_p] = []
        if runner is None:
            self._lib = _lib.library()
            if self._lib.pam_device_count() <= device:
                raise _lib.PamLibraryError(f"no HIP device {device}")
            for _ in range(slots):
                h = C.c_void_p()
                if self._lib.pam_context_create(C.byref(h), device) != 0 or not h.value:
                    raise _lib.PamLibraryError("pam_context_create failed")
                self._lib.pam_context_set_workspace_limit(h, self.optics_limit)
                self._ctx.append(h)
                if tuning:
                    from . import autotune
                    autotune.apply(self._lib, h, tuning)
        # workspace each slot's context holds (grow-only until trimmed); the
        # library rounds its allocations up by an eighth
        self._held = [0] * slots
        self.peak_reserved = 0

    def close(self):
        if self._lib is not None:
            for h in self._ctx:
                self._lib.pam_context_destroy(h)
        self._ctx = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- sizing and the device call ------------------------------------
    def need_bytes(self, req: Request) -> int:
        if self._need is not None:
            return int(self._need(req))
        return workspace_bytes(self._lib, req.pam, len(req.freqs), self.optics_limit)

    def _call(self, slot: int, req: Request):
        if self._runner is not None:
            return self._runner(slot, req)
        return _run_on_context(self._lib, self._ctx[slot], req)

    def _trim(self, slot: int):
        if self._lib is not None:
            self._lib.pam_context_trim(self._ctx[slot])
        self._held[slot] = 0

    def _admit(self, free: List[int], need: int) -> Optional[int]:
        """A free slot whose workspace can grow to `need` within the budget
        (idle slots give theirs back first), or None."""
        slot = max(free, key=lambda s: self._held[s])  # reuse warm workspace

        def total():
            return sum(self._held) - self._held[slot] + max(self._held[slot], need)

        for s in sorted(free, key=lambda s: -self._held[s]):
            if total() <= self.budget:
                break
            if s != slot and self._held[s]:
                self._trim(s)
        if total() > self.budget and self._held[slot] > need:
            self._trim(slot)
        if total() > self.budget:
            return None
        self._held[slot] = max(self._held[slot], need)
        self.peak_reserved = max(self.peak_reserved, sum(self._held))
        return slot

    # -- the loop ---------------------------------------------------------
    def run(self, requests: Sequence[Request], realtime: bool = True) -> List[Result]:
        """Serves `requests` (arrival times in seconds from now when
        `realtime`, else all at once in order) and returns their results in
        request order."""
        for r in requests:
            if len(getattr(getattr(r.pam, "df", None), "dataFullSpec", ())):
                # the workspace bound counts layers from hydro_q, which the
                # full-spectrum input does not use
                raise ValueError(f"request {r.rid}: full-spectrum input (hydro_fullspec) is "
                                 "not served; call runPamtra for it")
        order = sorted(range(len(requests)), key=lambda i: (requests[i].arrival, i))
        needs = [self.need_bytes(r) for r in requests]
        for i, n in enumerate(needs):
            if n * 9 // 8 > self.budget:
                raise ValueError(f"request {requests[i].rid} needs {n} B, over the "
                                 f"HBM budget of {self.budget} B")
        done: "queue.Queue" = queue.Queue()
        jobs = [queue.Queue() for _ in range(self.slots)]
        t0 = time.perf_counter()
        now = (lambda: time.perf_counter() - t0) if realtime else (lambda: 0.0)

        def worker(slot):
            while True:
                item = jobs[slot].get()
                if item is None:
                    return
                i, t_admit = item
                try:
                    st, msg, tb = self._call(slot, requests[i])
                except Exception as e:  # surfaced on the host, the loop goes on
                    st, msg, tb = _capi.PAM_FATAL, f"{type(e).__name__}: {e}", None
                done.put((i, slot, st, msg, tb, t_admit, now()))

        threads = [threading.Thread(target=worker, args=(s,), daemon=True)
                   for s in range(self.slots)]
        for th in threads:
            th.start()
        free = list(range(self.slots))
        busy = {}
        results: List[Optional[Result]] = [None] * len(requests)
        nxt = 0
        try:
            while nxt < len(order) or busy:
                # admit in arrival order while a slot and memory are free
                blocked = False
                while nxt < len(order) and free:
                    i = order[nxt]
                    if realtime and requests[i].arrival > now():
                        break
                    slot = self._admit(free, needs[i] * 9 // 8)
                    if slot is None:
                        blocked = True  # wait for a running request to finish
                        break
                    free.remove(slot)
                    busy[slot] = i
                    jobs[slot].put((i, now()))
                    nxt += 1
                # wait for a completion, or for the next arrival
                timeout = None
                if nxt < len(order) and free and not blocked and realtime:
                    timeout = max(0.0, requests[order[nxt]].arrival - now())
                try:
                    i, slot, st, msg, tb, t_admit, t_end = done.get(timeout=timeout)
                except queue.Empty:
                    continue
                r = requests[i]
                results[i] = Result(r.rid, st, msg, tb, r.arrival if realtime else 0.0,
                                    t_admit, t_end, slot, needs[i])
                if self._lib is not None:  # what the slot's context really holds now
                    self._held[slot] = int(self._lib.pam_context_bytes(self._ctx[slot]))
                del busy[slot]
                free.append(slot)
        finally:
            for q in jobs:
                q.put(None)
            for th in threads:
                th.join()
        return results  # type: ignore[return-value]


def summarize(results: Sequence[Result], requests: Sequence[Request], wall: float) -> dict:
    """Throughput (columns*channels/s over the wall time) and latency
    percentiles (ms) of a served set."""
    lat = np.array([r.latency for r in results]) * 1e3
    work = sum(q.columns * len(q.freqs) for q in requests)
    return {"requests": len(results),
            "columns_channels": int(work),
            "wall_s": round(wall, 4),
            "throughput": round(work / wall, 2) if wall > 0 else None,
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "failed": int(sum(1 for r in results if r.status >= _capi.PAM_FATAL))}


# ---------------------------------------------------------------------------
# Continuous batching
# ---------------------------------------------------------------------------

LEVEL_KEYS = ("hgt_lev", "temp_lev", "press_lev", "relhum_lev")
LAYER_KEYS = ("hgt", "temp", "press", "relhum", "hydro_q", "hydro_n", "hydro_reff")
COLUMN_KEYS = ("nlyrs", "groundtemp", "wind10u", "wind10v", "sfc_salinity", "sfc_emissivity",
               "sfc_type", "sfc_model", "sfc_refl")
MISSING = -9999.0


def _group_key(pam) -> tuple:
    """Requests whose columns may share one call: the same descriptor table,
    4D descriptor fields, settings and output levels."""
    sets = {k: v for k, v in pam.set.items() if not k.startswith("_") or k == "_nangles"}
    nml = dict(pam.nmlSet)
    return (pam.df.data.tobytes(), tuple(sorted(pam.df.data4D)), repr(sorted(sets.items())),
            repr(sorted((k, str(v)) for k, v in nml.items())), int(pam.p["noutlevels"]),
            int(getattr(pam, "_nangles", 16)))


def merge_columns(pams: Sequence[object]) -> dict:
    """One profile dict (and 4D descriptor dict) holding the columns of
    several requests, in order, along x.  Layers are padded to the largest
    max_nlyrs with the missing value; each column keeps its own nlyrs, so the
    padding is never read as atmosphere."""
    ML = max(int(p.p["max_nlyrs"]) for p in pams)
    nout = int(pams[0].p["noutlevels"])
    nh = len(pams[0].df.data)
    parts: dict = {}
    parts4: dict = {}

    def flat(a, ncol):
        a = np.asarray(a)
        return a.reshape((ncol,) + a.shape[2:])

    for pam in pams:
        p = pam.p
        ncol = int(p["ngridx"]) * int(p["ngridy"])
        L = int(p["max_nlyrs"])
        for key in LEVEL_KEYS + LAYER_KEYS:
            if key not in p:
                continue
            a = flat(p[key], ncol).astype(np.float64)
            want = ML + 1 if key in LEVEL_KEYS else ML
            if a.shape[1] < want:
                pad = np.full((ncol, want - a.shape[1]) + a.shape[2:], MISSING)
                a = np.concatenate([a, pad], axis=1)
            parts.setdefault(key, []).append(a)
        for key in COLUMN_KEYS:
            if key in p:
                parts.setdefault(key, []).append(flat(p[key], ncol))
        if "obs_height" in p:
            parts.setdefault("obs_height", []).append(flat(p["obs_height"], ncol).astype(np.float64))
        for key, a in pam.df.data4D.items():
            a = flat(a, ncol)
            if a.shape[1] < ML:
                fill = 0 if key == "nbin" else MISSING
                a = np.concatenate([a, np.full((ncol, ML - L) + a.shape[2:], fill,
                                               dtype=a.dtype)], axis=1)
            parts4.setdefault(key, []).append(a)
    ntot = sum(x.shape[0] for x in parts["nlyrs"])
    out = {k: np.concatenate(v, axis=0).reshape((ntot, 1) + v[0].shape[1:])
           for k, v in parts.items()}
    out.update({"ngridx": ntot, "ngridy": 1, "max_nlyrs": ML, "noutlevels": nout, "nhydro": nh})
    d4 = {k: np.concatenate(v, axis=0).reshape((ntot, 1) + v[0].shape[1:])
          for k, v in parts4.items()}
    return out, d4


class BatchingServer:
    """Continuous batching of forward-operator requests on one device.

    Requests that are waiting when the device frees up are packed into
    shared calls instead of running one call each: requests with the same
    descriptor table, settings and output levels (``_group_key``) form a
    group; for each channel of the group one call runs over the columns of
    every request that asked for it (columns concatenated along x, layers
    padded), so small requests fill the launches of the others.  A request
    is returned as soon as the calls of all its channels are done.  ``workers`` contexts
    (each a host thread and stream) alternate, so one batch is assembled on
    the host while the previous one runs.

    Each (column, channel) is still computed on its own by the same kernels,
    so a request's result is bitwise that of running it alone.
    """

    def __init__(self, device: int = 0, workers: int = 2, hbm_budget_gb: float = 64.0,
                 max_batch_columns: int = 16384, tuning: Optional[dict] = None):
        if workers < 1:
            raise ValueError("workers must be >= 1")
        self.workers = workers
        self.budget = int(hbm_budget_gb * (1 << 30))
        self.max_batch_columns = max_batch_columns
        self._lib = _lib.library()
        if self._lib.pam_device_count() <= device:
            raise _lib.PamLibraryError(f"no HIP device {device}")
        self._ctx: List[C.c_void_p] = []
        for _ in range(workers):
            h = C.c_void_p()
            if self._lib.pam_context_create(C.byref(h), device) != 0 or not h.value:
                raise _lib.PamLibraryError("pam_context_create failed")
            # each worker's optics in chunks of at most its share of the budget
            self._lib.pam_context_set_workspace_limit(h, int(0.6 * self.budget / workers))
            if tuning:
                from . import autotune
                autotune.apply(self._lib, h, tuning)
            self._ctx.append(h)
        self.calls = 0
        self.peak_reserved = 0

    def close(self):
        for h in self._ctx:
            self._lib.pam_context_destroy(h)
        self._ctx = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _run_batch(self, slot: int, reqs: Sequence[Request], finished=None):
        """All calls of one batch on worker `slot`; returns per request
        (status, message, tb).  `finished(rid, status, message, tb)` is
        called as soon as the last channel of a request is done."""
        out = {}
        left = {r.rid: len(r.freqs) for r in reqs}
        tbs = {}
        status = {}
        msgs = {}
        for r in reqs:
            p = r.pam.p
            nx, ny, nout = int(p["ngridx"]), int(p["ngridy"]), int(p["noutlevels"])
            nang = 2 * int(getattr(r.pam, "_nangles", 16))
            tbs[r.rid] = np.full((nx, ny, nout, nang, len(r.freqs), 2), np.nan)
            status[r.rid] = _capi.PAM_SUCCESS
            msgs[r.rid] = ""
        groups: dict = {}
        for r in reqs:
            groups.setdefault(_group_key(r.pam), []).append(r)
        lib, ctx = self._lib, self._ctx[slot]
        for members in groups.values():
            chans = sorted({float(f) for r in members for f in r.freqs})
            for ch in chans:
                use = [r for r in members if ch in [float(f) for f in r.freqs]]
                prof, d4 = merge_columns([r.pam for r in use])
                first = use[0].pam
                nml = dict(first.nmlSet)
                nml["active"] = False
                m = _capi.marshal(nml, first.set, first.df.data, d4, prof)
                f, fp = _capi.freqs_array([ch])
                hr = _capi.alloc_results(m, 1)
                st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                                 C.byref(m.profile), fp, 1, C.byref(hr.struct))
                self.calls += 1
                self.peak_reserved = max(self.peak_reserved,
                                         sum(int(lib.pam_context_bytes(h)) for h in self._ctx))
                tb = hr.arrays["tb"]
                sts = hr.arrays.get("status")
                msg = _capi.results_message(hr)
                o = 0
                for r in use:
                    p = r.pam.p
                    nx, ny = int(p["ngridx"]), int(p["ngridy"])
                    n = nx * ny
                    k = [float(v) for v in r.freqs].index(ch)
                    tbs[r.rid][..., k, :] = tb[o:o + n, 0, ..., 0, :].reshape(
                        tbs[r.rid][..., k, :].shape)
                    rs = int(st) if sts is None else int(np.max(sts[o:o + n] & 0xff))
                    if rs > status[r.rid]:
                        status[r.rid], msgs[r.rid] = rs, msg
                    o += n
                    left[r.rid] -= 1
                    if left[r.rid] == 0 and finished is not None:
                        finished(r.rid, status[r.rid], msgs[r.rid], tbs[r.rid])
        for r in reqs:
            out[r.rid] = (status[r.rid], msgs[r.rid], tbs[r.rid])
        return out

    def run(self, requests: Sequence[Request], realtime: bool = True) -> List[Result]:
        """Serves `requests` (arrival times in seconds from now when
        `realtime`, else all at once) and returns their results in request
        order."""
        for r in requests:
            if len(getattr(getattr(r.pam, "df", None), "dataFullSpec", ())):
                raise ValueError(f"request {r.rid}: full-spectrum input (hydro_fullspec) is "
                                 "not served; call runPamtra for it")
        order = sorted(range(len(requests)), key=lambda i: (requests[i].arrival, i))
        by_rid = {requests[i].rid: i for i in range(len(requests))}
        if len(by_rid) != len(requests):
            raise ValueError("request ids must be unique")
        done: "queue.Queue" = queue.Queue()
        jobs = [queue.Queue() for _ in range(self.workers)]
        t0 = time.perf_counter()
        now = (lambda: time.perf_counter() - t0) if realtime else (lambda: 0.0)

        def worker(slot):
            while True:
                item = jobs[slot].get()
                if item is None:
                    return
                batch, t_start = item
                reported = set()

                def finished(rid, st, msg, tb):
                    reported.add(rid)
                    done.put(("req", by_rid[rid], slot, (st, msg, tb), t_start, now()))

                try:
                    self._run_batch(slot, [requests[i] for i in batch], finished)
                except Exception as e:  # surfaced per request, the loop goes on
                    for i in batch:
                        if requests[i].rid not in reported:
                            done.put(("req", i, slot, (_capi.PAM_FATAL,
                                                       f"{type(e).__name__}: {e}", None),
                                      t_start, now()))
                done.put(("batch", None, slot, None, t_start, now()))

        threads = [threading.Thread(target=worker, args=(s,), daemon=True)
                   for s in range(self.workers)]
        for th in threads:
            th.start()
        free = list(range(self.workers))
        results: List[Optional[Result]] = [None] * len(requests)
        nxt = 0
        running = 0
        try:
            while nxt < len(order) or running:
                # a free worker takes every request that has arrived, up to
                # the batch's column limit (at least one)
                while free and nxt < len(order) and (not realtime or
                                                     requests[order[nxt]].arrival <= now()):
                    batch, cols = [], 0
                    while nxt < len(order) and (not realtime or
                                                requests[order[nxt]].arrival <= now()):
                        i = order[nxt]
                        c = requests[i].columns
                        if batch and cols + c > self.max_batch_columns:
                            break
                        batch.append(i)
                        cols += c
                        nxt += 1
                    slot = free.pop(0)
                    jobs[slot].put((batch, now()))
                    running += 1
                timeout = None
                if nxt < len(order) and free and realtime:
                    timeout = max(0.0, requests[order[nxt]].arrival - now())
                try:
                    kind, i, slot, res, t_start, t_end = done.get(timeout=timeout)
                except queue.Empty:
                    continue
                if kind == "req":
                    r = requests[i]
                    st, msg, tb = res
                    results[i] = Result(r.rid, st, msg, tb, r.arrival if realtime else 0.0,
                                        t_start, t_end, slot, 0)
                else:
                    running -= 1
                    free.append(slot)
        finally:
            for q in jobs:
                q.put(None)
            for th in threads:
                th.join()
        return results  # type: ignore[return-value]
