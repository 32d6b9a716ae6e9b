"""Benchmark of the passive forward operator on one MI355X.

Workload (default): the batched FP64 configuration -- 16384 columns with
20-120 layers (tops 10-20 km), 7 channels (22.235-183.31 GHz), cloud
(modified gamma) + rain (Marshall-Palmer) with 80 Mie bins each, NSTOKES=2,
16 Lobatto angles, ocean (TESSEM2) or Lambertian land.  Synthetic seeded
profiles; inputs resident in HBM before the timed region.  A step is one
pam_run_device call over the whole batch (all columns x channels).

Prints one JSON line (metric, value in columns*channels/s, roofline of the
dominant kernel from its HIP-event stage time, CPU baseline of the oracle on
a bounded sample of the same columns with its thread count).
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "columns·channels/s at ≤0.05 K TB error vs FP64; p50/p99 request latency"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector)
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0
NN = 32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def slice_columns(pam, n):
    """The first n columns (along x) of a pyPamtra profile dict."""
    p = {}
    nx = pam.p["ngridx"]
    for k, v in pam.p.items():
        if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] == nx:
            p[k] = v[:n]
        else:
            p[k] = v
    p["ngridx"] = n
    return p


def to_device(pam, torch):
    from pamtra_amd import _capi
    prof = {}
    for k, v in pam.p.items():
        if k in _capi.PROFILE_DOUBLE_KEYS or k == "obs_height":
            prof[k] = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device="cuda")
        elif k in ("nlyrs", "sfc_type", "sfc_model"):
            prof[k] = torch.as_tensor(np.ascontiguousarray(v).astype(np.int32), device="cuda")
        elif k == "sfc_refl":
            b = np.frombuffer("".join(np.asarray(v).reshape(-1)).encode(), dtype=np.uint8)
            prof[k] = torch.as_tensor(b.copy(), device="cuda")
        else:
            prof[k] = v
    return prof


def to_device_4d(pam, torch, ncol=None):
    """The 4D descriptor arrays (nx, ny, L, nhydro) in HBM: pam_run_device
    reads every array from the device."""
    out = {}
    for k, v in pam.df.data4D.items():
        a = np.ascontiguousarray(v, dtype=np.int32 if k == "nbin" else np.float64)
        out[k] = torch.as_tensor(a if ncol is None else a[:ncol].copy(), device="cuda")
    return out


# HBM bytes per launch of the dominant kernel on the default workload, from the
# two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) scaled by a copy of known
# size; recorded in this summary with the commands that produced it.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "batched_fp64_16384col_7ch_hbm_counters.md")


def measured_traffic(kernel):
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                f = line.split()
                if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
                    return int(f[2])
    except OSError:
        pass
    return None


_BATCHED_RATE = {}  # precision -> col*ch/s of the serving mix served all at once


def serve(args, precision="fp64", tuning=None, rate=50.0):
    """The serving mix at an offered Poisson rate: throughput and p50/p99
    latency (arrival to result on the host, input staging included)."""
    import numpy as np
    from pamtra_amd import serving, workloads
    rng = np.random.default_rng(1004)
    t0 = time.time()
    mix = workloads.serving_requests(args.serve_requests, seed=1004)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(mix)))
    for p, _ in mix:
        p.set["precision"] = precision
    reqs = [serving.Request(pam=p, freqs=f, arrival=float(a), rid=i)
            for i, ((p, f), a) in enumerate(zip(mix, arrivals))]
    log(f"serving mix built in {time.time() - t0:.1f} s: {len(reqs)} requests")
    shapes = None
    if tuning:
        # persistent grids only: a graph is not replayed across requests of
        # different shapes
        shapes = {k: v for k, v in tuning.items() if not k.startswith("_") and k != "graph"}
    if args.serve_mode == "batching":
        srv = serving.BatchingServer(workers=2, hbm_budget_gb=64.0, tuning=shapes)
    else:
        srv = serving.ForwardServer(slots=args.serve_slots, hbm_budget_gb=64.0, tuning=shapes)
    with srv:
        srv.run(reqs[:2], realtime=False)  # warm the workspace and code
        t0 = time.perf_counter()
        res = srv.run(reqs, realtime=True)
        wall = time.perf_counter() - t0
        peak = srv.peak_reserved
        calls = getattr(srv, "calls", None)
        work = sum(r.columns * len(r.freqs) for r in reqs)
        if precision not in _BATCHED_RATE:
            # the same requests all waiting at once: the batched rate of this mix
            t0 = time.perf_counter()
            srv.run(reqs, realtime=False)
            _BATCHED_RATE[precision] = work / (time.perf_counter() - t0)
    summ = serving.summarize(res, reqs, wall)
    summ["batched_rate_same_mix"] = round(_BATCHED_RATE[precision], 2)
    summ["fraction_of_batched"] = round(summ["throughput"] / _BATCHED_RATE[precision], 3)
    summ.update({"offered_req_s": rate, "mode": args.serve_mode,
                 "slots": 2 if args.serve_mode == "batching" else args.serve_slots,
                 "hbm_budget_gib": 64.0, "peak_workspace_gib": round(peak / 2 ** 30, 3),
                 "precision": precision, "launch_tuning": shapes, "device_calls": calls,
                 "mix": "1-4096 columns (log-uniform), 20-150 layers, 1-12 channels of "
                        "10.65-325 GHz, 0-4 of cloud/ice/snow/graupel with 100 bins"})
    log(f"serving: {summ}")
    return summ


def plain_dropin_rate(args):
    """The plain FP64 drop-in on the serving mix: each request through
    pyPamtra.runPamtra in turn, as a user of the reference would call it
    (host arrays staged over PCIe, one call per request, no batching across
    requests).  The serving legs are compared with this rate."""
    from pamtra_amd import workloads
    mix = workloads.serving_requests(args.serve_requests, seed=1004)
    # checkData=False: the padded levels above a column's own layers hold
    # values the range check refuses (the serving path reads only the column's
    # layers and skips the host check too)
    mix[0][0].runPamtra(mix[0][1], checkData=False)  # warm the context and the tables
    work = 0
    t0 = time.perf_counter()
    for i, (p, f) in enumerate(mix):
        p.set["precision"] = "fp64"
        p.runPamtra(f, checkData=False)
        work += p.p["ngridx"] * p.p["ngridy"] * len(f)
        if i % 20 == 19:
            log(f"plain drop-in: {i + 1} requests done")
    dt = time.perf_counter() - t0
    rate = work / dt
    log(f"plain FP64 drop-in on the serving mix: {rate:.0f} col*ch/s ({dt:.1f} s)")
    return {"value": round(rate, 2), "seconds": round(dt, 3), "columns_channels": int(work),
            "path": "pyPamtra.runPamtra per request, FP64, host arrays, one request at a time"}


def large_domain(args, torch):
    """The large-domain configuration (150 layers, 12 channels 10.65-325 GHz,
    four categories of 100 bins, 32 angles per hemisphere: 64x64 operators)
    on a column sample, FP64 and mixed; the full 1024 x 1024 domain streams
    the same calls in HBM-admitted column chunks."""
    from pamtra_amd import _capi, _lib, workloads
    t0 = time.time()
    pam, freqs = workloads.config5(ncol=args.large_cols, seed=1005)
    log(f"large-domain sample built in {time.time() - t0:.1f} s: {args.large_cols} columns")
    lib = _lib.library(32)
    ctx = _lib.get_context(0, 32)
    prof = to_device(pam, torch)
    d4 = to_device_4d(pam, torch)
    nml = dict(pam.nmlSet)
    nf = len(freqs)
    f, fp = _capi.freqs_array(freqs)
    res = {}
    tbs = {}
    for precision in ("fp64", "mixed"):
        sets = dict(pam.set, _nangles=32, precision=precision)
        m = _capi.marshal(nml, sets, pam.df.data, d4, prof)
        nx, ny, nout = m.nx, m.ny, m.noutlevels
        tb = torch.zeros((nx, ny, nout, 64, nf, 2), dtype=torch.float64, device="cuda")
        status = torch.zeros((nx, ny, nf), dtype=torch.int32, device="cuda")
        r = _capi.PamResults()
        r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
        r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))
        msg = C.create_string_buffer(512)
        r.msg = C.cast(msg, C.c_char_p)
        r.msglen = 512

        def call():
            if lib.pam_run_device(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                                  C.byref(m.profile), fp, nf, C.byref(r), None) == _capi.PAM_FATAL:
                raise RuntimeError("pam_run_device (32 angles): " + msg.value.decode())

        call()
        torch.cuda.synchronize()
        names = lib.pam_stage_names().decode().split(",")
        acc = np.zeros(len(names))
        t0 = time.perf_counter()
        for i in range(args.large_steps):
            call()
            ms = (C.c_double * len(names))()
            lib.pam_last_stage_times(ctx, ms, len(names))
            acc += np.array(ms[:])
            log(f"large-domain {precision} step {i} done")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.large_steps
        tbs[precision] = tb.cpu().numpy()
        work = (C.c_int64 * 14)()
        lib.pam_last_work(ctx, work, 14)
        stage = {n: v / args.large_steps for n, v in zip(names, acc)}
        res[precision] = {"value": round(nx * ny * nf / dt, 2), "ms_per_step": round(dt * 1e3, 3),
                          "stage_ms": {n: round(v, 3) for n, v in stage.items()},
                          "status_nonzero": int((status.cpu().numpy() != 0).sum()),
                          "doubling_steps": work[3], "doubling_products": work[9],
                          "adding_products": work[10]}
        if precision == "fp64":
            # the reference's algorithm: 14 N^3 flops per doubling step, N = 64
            ach = work[3] * 14 * 64 ** 3 / (stage["doubling"] * 1e-3) / 1e12
            res[precision]["doubling_roofline"] = {
                "achieved": round(ach, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 4),
                "mfma_executed_tflops": round(work[9] * 2 * 64 ** 3 / (stage["doubling"] * 1e-3)
                                              / 1e12, 3)}
        else:
            res[precision]["doubling_fp32_roofline"] = fp32_doubling_roofline(
                list(work), stage["doubling"], 64)
    dch = np.max(np.abs(tbs["mixed"] - tbs["fp64"]), axis=(0, 1, 2, 3, 5))
    res["mixed"]["max_abs_dtb_vs_fp64_K"] = {f"{fr:g}": float(v) for fr, v in zip(freqs, dch)}
    res["mixed"]["speedup_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    res["workload"] = (f"large domain sample: {args.large_cols} of the 1024x1024 columns, 150 "
                       "layers, 12 channels 10.65-325 GHz, cloud/ice/snow/graupel 100 bins, "
                       "NSTOKES=2, 32 angles (64x64 operators)")
    res["full_domain_s_fp64"] = round(1024 * 1024 * nf / res["fp64"]["value"], 1)
    log(f"large domain: {res}")
    return res


def large_domain_stream(args, torch):
    """The large-domain configuration at its full size, streamed: the 1024 x
    1024 columns are generated and run in slices of --large-slice columns,
    one pam_run_device call per slice with the slice's inputs already in HBM
    (generation and the host-to-device copies are outside the timed calls).
    The mixed path runs every slice; FP64 runs the first --large-fp64-slices
    slices, on which the per-channel TB difference is taken."""
    from pamtra_amd import _capi, _lib, workloads
    lib = _lib.library(32)
    ctx = _lib.get_context(0, 32)
    names = lib.pam_stage_names().decode().split(",")
    total, size = args.large_total, args.large_slice
    nslice = (total + size - 1) // size
    res = {p: {"seconds": 0.0, "columns": 0, "stage_ms": np.zeros(len(names)),
               "status_nonzero": 0} for p in ("fp64", "mixed")}
    dmax = None
    for k in range(nslice):
        n = min(size, total - k * size)
        pam, freqs = workloads.config5(ncol=n, seed=1005 + 7919 * k)
        prof = to_device(pam, torch)
        d4 = to_device_4d(pam, torch)
        nf = len(freqs)
        f, fp = _capi.freqs_array(freqs)
        tbs = {}
        for precision in (("fp64", "mixed") if k < args.large_fp64_slices else ("mixed",)):
            m = _capi.marshal(dict(pam.nmlSet), dict(pam.set, _nangles=32, precision=precision),
                              pam.df.data, d4, prof)
            tb = torch.zeros((m.nx, m.ny, m.noutlevels, 64, nf, 2), dtype=torch.float64,
                             device="cuda")
            status = torch.zeros((m.nx, m.ny, nf), dtype=torch.int32, device="cuda")
            r = _capi.PamResults()
            r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
            r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))
            msg = C.create_string_buffer(512)
            r.msg = C.cast(msg, C.c_char_p)
            r.msglen = 512
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if lib.pam_run_device(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                                  C.byref(m.profile), fp, nf, C.byref(r), None) == _capi.PAM_FATAL:
                raise RuntimeError("pam_run_device (large domain): " + msg.value.decode())
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            ms = (C.c_double * len(names))()
            lib.pam_last_stage_times(ctx, ms, len(names))
            q = res[precision]
            q["seconds"] += dt
            q["columns"] += n
            q["stage_ms"] += np.array(ms[:])
            q["status_nonzero"] += int((status.cpu().numpy() != 0).sum())
            if k < args.large_fp64_slices:
                tbs[precision] = tb.cpu().numpy()
            log(f"large domain slice {k + 1}/{nslice} ({n} columns) {precision}: {dt:.2f} s")
        if len(tbs) == 2:
            d = np.max(np.abs(tbs["mixed"] - tbs["fp64"]), axis=(0, 1, 2, 3, 5))
            dmax = d if dmax is None else np.maximum(dmax, d)
        del prof, d4
    out = {}
    for p, q in res.items():
        if q["columns"] == 0:
            continue
        out[p] = {"columns": q["columns"], "channels": nf, "seconds": round(q["seconds"], 2),
                  "value": round(q["columns"] * nf / q["seconds"], 2),
                  "stage_ms": {n: round(v, 1) for n, v in zip(names, q["stage_ms"])},
                  "status_nonzero": q["status_nonzero"]}
    out["mixed"]["max_abs_dtb_vs_fp64_K"] = {f"{fr:g}": float(v) for fr, v in zip(freqs, dmax)}
    out["mixed"]["speedup_vs_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    out["workload"] = (f"large domain, {total} columns generated and run in slices of {size} "
                       "(one call each), 150 layers, 12 channels 10.65-325 GHz, "
                       "cloud/ice/snow/graupel 100 bins, NSTOKES=2, 32 angles; FP64 on the first "
                       f"{args.large_fp64_slices} slices")
    log(f"large domain (streamed): {out}")
    return out


def fp32_doubling_roofline(work, dbl_ms, n):
    """The mixed doubling on the FP32 matrix cores: the N x N products it
    executed (work[9]: start layer, series inverses and doublings), 2 N^3
    flops each, against the 157.3 TFLOP/s FP32 matrix peak; beside it the
    reference's count, 14 N^3 per doubling step of the path's own steps."""
    ex = work[9] * 2 * n ** 3 / (dbl_ms * 1e-3) / 1e12
    ref = work[3] * 14 * n ** 3 / (dbl_ms * 1e-3) / 1e12
    return {"achieved": round(ex, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(ex / PEAK_FP32_TFLOPS, 4), "counted": "executed products",
            "reference_count_tflops": round(ref, 3)}


def mie_flops(work):
    """The Mie stage's flops counted on the reference's algorithm
    (mie_scat_utilities.f90:8-160, mie_spheres.f90:178-314): per bin the D_n
    recursion from nt+15 (20 flops a step), 76 flops per term for a_n, b_n
    and the efficiencies, 28 per term and angle for S1/S2 and the phase sums;
    per (layer, hydrometeor) 13 per angle and order for the Legendre
    projection.  Counters: [6] bins, [7] sum of series lengths, [12] sum of
    nquad * series length, [13] sum of nquad * (nlegen + 1)."""
    return 20 * (work[7] + 15 * work[6]) + 76 * work[7] + 28 * work[12] + 13 * work[13]


def config3_leg(args, torch):
    """The mixed-precision configuration at full size: 16384 columns, four
    categories (cloud, ice, snow, graupel) of 100 bins, config 2's grid and 7
    channels; FP64 and mixed, each timed, with the per-channel TB difference."""
    from pamtra_amd import _capi, _lib, workloads
    t0 = time.time()
    pam, freqs = workloads.config3(ncol=args.config3_cols, seed=1003)
    log(f"config-3 workload built in {time.time() - t0:.1f} s")
    lib, ctx = _lib.library(), _lib.get_context(0)
    prof = to_device(pam, torch)
    d4 = to_device_4d(pam, torch)
    nml = dict(pam.nmlSet)
    nf = len(freqs)
    f, fp = _capi.freqs_array(freqs)
    names = lib.pam_stage_names().decode().split(",")
    res, tbs = {}, {}
    for precision in ("fp64", "mixed"):
        m = _capi.marshal(nml, dict(pam.set, precision=precision), pam.df.data, d4, prof)
        nx, ny, nout = m.nx, m.ny, m.noutlevels
        tb = torch.zeros((nx, ny, nout, 32, nf, 2), dtype=torch.float64, device="cuda")
        status = torch.zeros((nx, ny, nf), dtype=torch.int32, device="cuda")
        r = _capi.PamResults()
        r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
        r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))

        def call():
            if lib.pam_run_device(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                                  C.byref(m.profile), fp, nf, C.byref(r), None) == _capi.PAM_FATAL:
                raise RuntimeError("pam_run_device (config 3) failed")

        call()
        torch.cuda.synchronize()
        acc = np.zeros(len(names))
        t0 = time.perf_counter()
        for i in range(args.config3_steps):
            call()
            ms = (C.c_double * len(names))()
            lib.pam_last_stage_times(ctx, ms, len(names))
            acc += np.array(ms[:])
            log(f"config-3 {precision} step {i} done")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.config3_steps
        tbs[precision] = tb.cpu().numpy()
        work = (C.c_int64 * 14)()
        lib.pam_last_work(ctx, work, 14)
        res[precision] = {"value": round(nx * ny * nf / dt, 2), "ms_per_step": round(dt * 1e3, 3),
                          "stage_ms": {n: round(v / args.config3_steps, 3)
                                       for n, v in zip(names, acc)},
                          "status_nonzero": int((status.cpu().numpy() != 0).sum())}
        if precision == "mixed":
            res[precision]["doubling_fp32_roofline"] = fp32_doubling_roofline(
                work, res[precision]["stage_ms"]["doubling"], NN)
    dch = np.max(np.abs(tbs["mixed"] - tbs["fp64"]), axis=(0, 1, 2, 3, 5))
    res["mixed"]["max_abs_dtb_vs_fp64_K"] = {f"{fr:g}": float(v) for fr, v in zip(freqs, dch)}
    res["mixed"]["speedup_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    res["workload"] = (f"{args.config3_cols} columns, 20-120 layers, 7 channels 22.235-183.31 "
                       "GHz, cloud/ice/snow/graupel Mie 100 bins each (snow soft spheres by "
                       "column), NSTOKES=2, 16 angles")
    log(f"config 3: {res}")
    return res


def tmatrix_leg(args, torch):
    """The T-matrix producer: tmatrix_config (the config-1 grid, cloud
    droplets as T-matrix spheres, rain as oblate spheroids of axis ratio
    0.7, 30 bins each, 3 channels), FP64, timed per call, with the solve
    count and rate, and the oracle (binary128 restatement) on a few columns
    as the CPU baseline and parity check."""
    from pamtra_amd import _capi, _lib, workloads
    pam, freqs = workloads.tmatrix_config(ncol=args.tmatrix_cols)
    lib, ctx = _lib.library(), _lib.get_context(0)
    prof = to_device(pam, torch)
    nml = dict(pam.nmlSet)
    nf = len(freqs)
    f, fp = _capi.freqs_array(freqs)
    names = lib.pam_stage_names().decode().split(",")
    m = _capi.marshal(nml, pam.set, pam.df.data, pam.df.data4D, prof)
    nx, ny, nout = m.nx, m.ny, m.noutlevels
    tb = torch.zeros((nx, ny, nout, 32, nf, 2), dtype=torch.float64, device="cuda")
    status = torch.zeros((nx, ny, nf), dtype=torch.int32, device="cuda")
    r = _capi.PamResults()
    r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
    r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))

    def call():
        if lib.pam_run_device(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                              C.byref(m.profile), fp, nf, C.byref(r), None) == _capi.PAM_FATAL:
            raise RuntimeError("pam_run_device (T-matrix) failed")

    call()
    torch.cuda.synchronize()
    acc = np.zeros(len(names))
    t0 = time.perf_counter()
    for i in range(args.tmatrix_steps):
        call()
        ms = (C.c_double * len(names))()
        lib.pam_last_stage_times(ctx, ms, len(names))
        acc += np.array(ms[:])
        log(f"tmatrix step {i} done")
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.tmatrix_steps
    work = (C.c_int64 * 14)()
    lib.pam_last_work(ctx, work, 14)
    stage = {n: round(v / args.tmatrix_steps, 3) for n, v in zip(names, acc)}
    particles = int(work[6])
    res = {"value": round(nx * ny * nf / dt, 2), "unit": "columns·channels/s",
           "ms_per_step": round(dt * 1e3, 3), "stage_ms": stage,
           "producer_stage": "mie", "particles_per_call": particles,
           "mean_tmatrix_order": round(work[7] / max(1, particles), 2),
           "particles_per_s": round(particles / (stage["mie"] * 1e-3), 1),
           "status_nonzero": int((status.cpu().numpy() != 0).sum()),
           "workload": (f"{args.tmatrix_cols} columns x 30 layers (0-12 km, ocean), 3 channels "
                        "22.235/35.5/89 GHz, cloud (T-matrix spheres) and rain (oblate "
                        "spheroids, as_ratio 0.7), 30 bins each, canting 0, FP64")}
    if args.tmatrix_cpu_cols > 0 and not args.no_cpu:
        import oracle
        n = min(args.tmatrix_cpu_cols, nx)
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (
            os.cpu_count() or 1)
        threads = args.cpu_threads or min(16, avail)
        sub = slice_columns(pam, n)
        t1 = time.perf_counter()
        ref, st, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, sub, freqs,
                                nthreads=threads)
        tc = time.perf_counter() - t1
        tbh = tb.cpu().numpy()
        ok = np.isfinite(ref["tb"])
        res["max_abs_dtb_vs_oracle_K"] = float(np.max(np.abs(tbh[:n][ok] - ref["tb"][ok])))
        res["cpu_baseline"] = {"value": round(n * nf / tc, 4), "unit": "columns·channels/s",
                               "cores": threads, "kind": "port",
                               "sample": f"first {n} of the {nx} columns x {nf} channels, "
                                         f"{tc:.1f} s (binary128 T-matrix solves)"}
    log(f"tmatrix: {res}")
    return res


def aggregates(args):
    """Cost of the self-similar Rayleigh-Gans producer: the four-category
    columns (config 3 shapes) with snow and graupel as Mie soft spheres and
    as SSRGA aggregates, FP64, stage times of each."""
    from pamtra_amd import _capi, _lib, workloads
    lib, ctx = _lib.library(), _lib.get_context(0)
    res = {}
    for kind in ("mie", "ssrg"):
        pam, freqs = workloads.config3(ncol=args.aggregates_cols, seed=1003)
        if kind == "ssrg":
            pam.df.data["scat_name"][2] = "ssrg-rt3_0.19_0.23_1.67"
            pam.df.data["as_ratio"][2] = 0.6
            pam.df.data["scat_name"][3] = "ssrg-rt3"
            pam.df.data["as_ratio"][3] = 0.8
        m = _capi.marshal(dict(pam.nmlSet), pam.set, pam.df.data, pam.df.data4D, pam.p)
        f, fp = _capi.freqs_array(freqs)
        hr = _capi.alloc_results(m, len(f))
        names = lib.pam_stage_names().decode().split(",")
        acc = np.zeros(len(names))
        for i in range(3):
            st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                             C.byref(m.profile), fp, len(f), C.byref(hr.struct))
            if st == _capi.PAM_FATAL:
                raise RuntimeError("aggregates: " + _capi.results_message(hr))
            ms = (C.c_double * len(names))()
            lib.pam_last_stage_times(ctx, ms, len(names))
            if i:
                acc += np.array(ms[:])
            log(f"aggregates {kind} call {i} done")
        res[kind] = {n: round(v / 2, 3) for n, v in zip(names, acc)}
    res["workload"] = (f"{args.aggregates_cols} columns of the four-category config-3 shapes, "
                       "7 channels; snow and graupel as Mie soft spheres or SSRGA aggregates")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncol", type=int, default=16384)
    ap.add_argument("--cpu-cols", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run")
    ap.add_argument("--no-host-entry", action="store_true",
                    help="skip the pam_run (host arrays, PCIe-inclusive) measurement")
    ap.add_argument("--serve-requests", type=int, default=100,
                    help="requests of the serving mix (0: no serving run)")
    ap.add_argument("--serve-rate", type=str, default="50,100,200,500,1000,2000",
                    help="offered requests/s (Poisson), comma-separated: one serving run each")
    ap.add_argument("--serve-mode", choices=("batching", "slots"), default="batching",
                    help="batching: waiting requests share calls per channel; slots: one "
                         "call per request on concurrent contexts")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving mix and print its summary")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--tune-cols", type=int, default=0,
                    help="columns of the workload slice the autotuner measures on "
                         "(0: the whole workload, so that every launch of a kernel has "
                         "the timed run's shape)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also fill and then sum a buffer of MIB MiB (kernels of known "
                         "HBM traffic, to scale the HBM counters)")
    ap.add_argument("--large-cols", type=int, default=2048,
                    help="columns of the large-domain (32-angle) sample (0: skip it)")
    ap.add_argument("--large-steps", type=int, default=2)
    ap.add_argument("--config3-cols", type=int, default=16384,
                    help="columns of the four-category mixed-precision configuration "
                         "(0: skip it)")
    ap.add_argument("--config3-steps", type=int, default=2)
    ap.add_argument("--aggregates-cols", type=int, default=0,
                    help="also time the SSRGA producer on this many four-category columns")
    ap.add_argument("--aggregates-only", action="store_true",
                    help="only the SSRGA producer timing (needs --aggregates-cols)")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the large-domain (32-angle) sample and print its summary")
    ap.add_argument("--large-stream", action="store_true",
                    help="run only the large domain at full size, streamed in column slices")
    ap.add_argument("--large-total", type=int, default=1024 * 1024)
    ap.add_argument("--large-slice", type=int, default=65536)
    ap.add_argument("--large-fp64-slices", type=int, default=2)
    ap.add_argument("--tmatrix-cols", type=int, default=1024,
                    help="columns of the T-matrix leg (0: skip it)")
    ap.add_argument("--tmatrix-steps", type=int, default=2)
    ap.add_argument("--tmatrix-cpu-cols", type=int, default=2,
                    help="columns of the T-matrix leg's oracle sample")
    ap.add_argument("--tmatrix-only", action="store_true",
                    help="run only the T-matrix leg and print its summary")
    ap.add_argument("--config3-only", action="store_true",
                    help="run only the four-category configuration (FP64 and mixed) and print "
                         "its summary")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the columns are "
              "independent, a multi-GPU run would shard them with no collective.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from pamtra_amd import _capi, _lib, workloads

    if args.large_only:
        print(json.dumps({"large_domain": large_domain(args, torch)}))
        return
    if args.large_stream:
        print(json.dumps({"large_domain_full": large_domain_stream(args, torch)}))
        return
    if args.config3_only:
        print(json.dumps({"config3": config3_leg(args, torch)}))
        return
    if args.tmatrix_only:
        print(json.dumps({"tmatrix": tmatrix_leg(args, torch)}))
        return
    if args.aggregates_only:
        print(json.dumps({"aggregates": aggregates(args)}))
        return
    rates = [float(x) for x in str(args.serve_rate).split(",") if x.strip()]
    if args.serve_only:
        plain = plain_dropin_rate(args)
        res = {"plain_fp64_dropin": plain, "fp64": [serve(args, rate=r) for r in rates],
               "mixed": [serve(args, "mixed", rate=r) for r in rates]}
        for p in ("fp64", "mixed"):
            for e in res[p]:
                e["vs_plain_fp64_dropin"] = round(e["throughput"] / plain["value"], 3)
        print(json.dumps({"serving": res}))
        return

    t0 = time.time()
    pam, freqs = workloads.config2(ncol=args.ncol, seed=1002)
    log(f"workload built in {time.time() - t0:.1f} s: {args.ncol} columns x {len(freqs)} ch")
    lib = _lib.library()
    ctx = _lib.get_context(0)
    prof = to_device(pam, torch)
    d4 = to_device_4d(pam, torch)
    nml = dict(pam.nmlSet)
    m = _capi.marshal(nml, pam.set, pam.df.data, d4, prof)
    nx, ny, nout, nf = m.nx, m.ny, m.noutlevels, len(freqs)
    tb = torch.zeros((nx, ny, nout, 32, nf, 2), dtype=torch.float64, device="cuda")
    status = torch.zeros((nx, ny, nf), dtype=torch.int32, device="cuda")
    r = _capi.PamResults()
    r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
    r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))
    msg = C.create_string_buffer(512)
    r.msg = C.cast(msg, C.c_char_p)
    r.msglen = 512
    f, fp = _capi.freqs_array(freqs)

    def step():
        st = lib.pam_run_device(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                                C.byref(m.profile), fp, nf, C.byref(r), None)
        if st == _capi.PAM_FATAL:
            raise RuntimeError("pam_run_device: " + msg.value.decode())
        return st

    def slice_runner(precision):
        """One synchronous call on the first --tune-cols columns (the
        autotuner's sample of the workload)."""
        ns = min(args.tune_cols, nx) if args.tune_cols > 0 else nx
        sub = type("Sub", (), {})()
        sub.p = slice_columns(pam, ns)
        ps = to_device(sub, torch)
        sets_s = dict(pam.set)
        sets_s["precision"] = precision
        ms_ = _capi.marshal(nml, sets_s, pam.df.data, to_device_4d(pam, torch, ns), ps)
        tb_s = torch.zeros((ns, ny, nout, 32, nf, 2), dtype=torch.float64, device="cuda")
        st_s = torch.zeros((ns, ny, nf), dtype=torch.int32, device="cuda")
        rs = _capi.PamResults()
        rs.tb = C.cast(C.c_void_p(tb_s.data_ptr()), C.POINTER(C.c_double))
        rs.status = C.cast(C.c_void_p(st_s.data_ptr()), C.POINTER(C.c_int32))
        keep = (ps, ms_, tb_s, st_s)

        def run_once():
            _ = keep
            if lib.pam_run_device(ctx, C.byref(ms_.settings), ms_.hydros, C.byref(ms_.hydro4d),
                                  C.byref(ms_.profile), fp, nf, C.byref(rs), None) == _capi.PAM_FATAL:
                raise RuntimeError("pam_run_device (tuning sample) failed")
            torch.cuda.synchronize()
        return run_once

    def timed(fn, steps):
        names_ = lib.pam_stage_names().decode().split(",")
        acc = np.zeros(len(names_))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            fn()
            ms = (C.c_double * len(names_))()
            lib.pam_last_stage_times(ctx, ms, len(names_))
            acc += np.array(ms[:])
            log(f"step {i} done")
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, acc / steps

    for i in range(args.warmup):
        step()
        log(f"warmup {i} done")
    torch.cuda.synchronize()
    names = lib.pam_stage_names().decode().split(",")
    nst = len(names)
    tuning = None
    if not args.no_tune:
        from pamtra_amd import autotune
        dt0, stage0 = timed(step, args.steps)
        log(f"default launch shapes: {dt0 * 1e3:.1f} ms/step")
        t0 = time.time()
        ch = autotune.tune(lib, ctx, slice_runner("fp64"))
        log(f"autotuned in {time.time() - t0:.1f} s: {ch}")
        tuning = {"fp64": ch, "default_shapes_value": round(nx * ny * nf / dt0, 2),
                  "default_shapes_stage_ms": {n: round(v, 3) for n, v in zip(names, stage0)},
                  "sample_columns": min(args.tune_cols, nx) if args.tune_cols > 0 else nx}
        for i in range(args.warmup):
            step()
    dt, stage_ms = timed(step, args.steps)
    work = (C.c_int64 * 14)()
    lib.pam_last_work(ctx, work, 14)
    work = list(work)
    value = nx * ny * nf / dt

    # Algorithmic flops of the RT stages, counted on the reference's algorithm
    # (one FMA = 2 flops): a symmetric doubling step is 6 products and one
    # inverse, 14 N^3 (radintg4.f90:212-297); the m=0 phase matrix is
    # 4 NUMMU^2 (NUMPTS/2+1)(nlegen+1) terms of 11 flops (scatcnv.f90:103-192);
    # adding with prefix reuse is 20 N^3 per layer combine plus 2 x 8 N^3 for
    # the two internal radiances (radintg4.f90:119-172, 447-513).
    n3 = NN ** 3
    nummu = NN // 2
    f_doubling = work[3] * 14 * n3
    f_phase = work[8] * 4 * nummu * nummu * 11
    f_adding = (work[4] + work[5] + nx * ny * nf) * 20 * n3 + nx * ny * nf * 16 * n3
    flops = {"doubling": f_doubling, "phase": f_phase, "adding": f_adding}
    # 32x32 products the matrix cores actually executed (series inverse included)
    executed = {"doubling": work[9] * 2 * n3, "adding": work[10] * 2 * n3}
    dom = max(range(nst), key=lambda i: stage_ms[i])
    dom_name = names[dom]
    roof = None
    if dom_name in flops:
        ach = flops[dom_name] / (stage_ms[dom] * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": dom_name, "achieved": round(ach, 3),
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP64_TFLOPS, 4),
                "traffic": (measured_traffic("k_" + dom_name)
                            if args.ncol == 16384 else None),
                "mfma_executed_tflops": (round(executed[dom_name] / (stage_ms[dom] * 1e-3) / 1e12, 3)
                                         if dom_name in executed else None)}

    # The Mie stage against the FP64 peak, counted on the reference's
    # algorithm (mie_flops)
    f_mie = mie_flops(work)
    mie_ms = stage_ms[names.index("mie")]
    ach_mie = f_mie / (mie_ms * 1e-3) / 1e12
    mie_roof = {"bound": "fp64", "kernel": "mie", "achieved": round(ach_mie, 3),
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach_mie / PEAK_FP64_TFLOPS, 4), "flops": f_mie,
                "bins": work[6], "series_terms": work[7]}

    tb_h = tb.cpu().numpy()
    st_h = status.cpu().numpy()

    # the mixed-precision path on the same workload: throughput and its
    # per-channel TB error against the FP64 result above
    mixed = None
    if not args.no_mixed:
        sets_m = dict(pam.set)
        sets_m["precision"] = "mixed"
        mm = _capi.marshal(nml, sets_m, pam.df.data, d4, prof)

        def step_m():
            st = lib.pam_run_device(ctx, C.byref(mm.settings), mm.hydros, C.byref(mm.hydro4d),
                                    C.byref(mm.profile), fp, nf, C.byref(r), None)
            if st == _capi.PAM_FATAL:
                raise RuntimeError("pam_run_device (mixed): " + msg.value.decode())

        if not args.no_tune:
            from pamtra_amd import autotune
            chm = autotune.tune(lib, ctx, slice_runner("mixed"))
            log(f"autotuned (mixed): {chm}")
            tuning["mixed"] = chm
        for _ in range(args.warmup):
            step_m()
        torch.cuda.synchronize()
        sms = np.zeros(nst)
        t0 = time.perf_counter()
        for i in range(args.steps):
            step_m()
            ms = (C.c_double * nst)()
            lib.pam_last_stage_times(ctx, ms, nst)
            sms += np.array(ms[:])
        torch.cuda.synchronize()
        dtm = (time.perf_counter() - t0) / args.steps
        workm = (C.c_int64 * 14)()
        lib.pam_last_work(ctx, workm, 14)
        workm = list(workm)
        tbm = tb.cpu().numpy()
        dch = np.max(np.abs(tbm - tb_h), axis=(0, 1, 2, 3, 5))
        sm = {n: round(v / args.steps, 3) for n, v in zip(names, sms)}
        f_miem = mie_flops(workm)
        mixed = {"value": round(nx * ny * nf / dtm, 2), "ms_per_step": round(dtm * 1e3, 3),
                 "stage_ms": sm,
                 "speedup_vs_fp64": round(dt / dtm, 3),
                 "max_abs_dtb_vs_fp64_K": {f"{fr:g}": float(v) for fr, v in zip(freqs, dch)},
                 "doubling_fp32_roofline": fp32_doubling_roofline(workm, sm["doubling"], NN),
                 "mie_roofline": {"bound": "fp32", "achieved": round(
                     f_miem / (sm["mie"] * 1e-3) / 1e12, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(
                         f_miem / (sm["mie"] * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                     "flops": f_miem},
                 "dtype": "f32: Mie series (D_n, a_n, b_n) and S1/S2 products, the doubling "
                          "(start layer included) and the adding; f64: psi/chi recursion, "
                          "efficiency and phase sums, Legendre projection, phase matrix, "
                          "CHECK_NORM4 and the choice of the doubling count"}
        log(f"mixed: {mixed}")

    host_rate = None
    if not args.no_host_entry:
        # the host-array entry: inputs staged to HBM and TB copied back inside the call
        mh = _capi.marshal(nml, pam.set, pam.df.data, pam.df.data4D, pam.p)
        hr = _capi.alloc_results(mh, nf)
        t0 = time.perf_counter()
        sth = lib.pam_run(ctx, C.byref(mh.settings), mh.hydros, C.byref(mh.hydro4d),
                          C.byref(mh.profile), fp, nf, C.byref(hr.struct))
        th = time.perf_counter() - t0
        if sth == _capi.PAM_FATAL:
            raise RuntimeError("pam_run: " + _capi.results_message(hr))
        host_rate = nx * ny * nf / th
        log(f"host-array entry (PCIe-inclusive): {host_rate:.0f} col*ch/s")
    if args.copy_calibration:
        # kernels of known traffic for the HBM counters: a fill writes the
        # buffer once, a sum reads it once
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        total = float(a.sum())
        torch.cuda.synchronize()
        log(f"counter calibration: {args.copy_calibration} MiB written (fill), "
            f"then read (sum = {total:.0f})")
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "columns·channels/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched FP64 drop-in: 16384 columns, 20-120 layers, tops "
                               "10-20 km, 7 channels 22.235-183.31 GHz, cloud+rain Mie 80 bins, "
                               "NSTOKES=2, 16 angles, ocean TESSEM2 / Lambertian land",
                   "columns": nx * ny, "channels": nf, "precision": "fp64"},
        "stage_ms": {n: round(v, 3) for n, v in zip(names, stage_ms)},
        "work": {"scattering_items": work[2], "doubling_steps": work[3],
                 "adding_scattering": work[4], "adding_clear": work[5],
                 "doubling_products": work[9], "adding_products": work[10],
                 "gauss_jordan_units": work[11]},
        "status_nonzero": int((st_h != 0).sum()),
        "pcie_inclusive_value": None if host_rate is None else round(host_rate, 2),
        "mixed_precision": mixed,
        "launch_tuning": tuning,
        "roofline": roof,
        "mie_roofline": mie_roof,
    }

    if args.serve_requests > 0:
        plain = plain_dropin_rate(args)
        out["serving"] = {"plain_fp64_dropin": plain,
                          "fp64": [serve(args, tuning=tuning and tuning.get("fp64"), rate=r)
                                   for r in rates]}
        if not args.no_mixed:
            out["serving"]["mixed"] = [serve(args, "mixed", tuning=tuning and tuning.get("mixed"),
                                             rate=r) for r in rates]
        for p in ("fp64", "mixed"):
            for e in out["serving"].get(p, []):
                e["vs_plain_fp64_dropin"] = round(e["throughput"] / plain["value"], 3)

    if args.config3_cols > 0:
        out["config3"] = config3_leg(args, torch)
    if args.large_cols > 0:
        out["large_domain"] = large_domain(args, torch)
    if args.aggregates_cols > 0:
        out["aggregates"] = aggregates(args)
    if args.tmatrix_cols > 0:
        out["tmatrix"] = tmatrix_leg(args, torch)

    if not args.no_cpu:
        import oracle
        n = min(args.cpu_cols, nx)
        # every host core this process may use, at most the 16 a one-GPU box
        # grants (its nproc can show more)
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (
            os.cpu_count() or 1)
        threads = args.cpu_threads or min(16, avail)
        sub = slice_columns(pam, n)
        t0 = time.perf_counter()
        ref, st, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, sub, freqs,
                                nthreads=threads)
        tc = time.perf_counter() - t0
        ok = np.isfinite(ref["tb"])
        dmax = float(np.max(np.abs(tb_h[:n][ok] - ref["tb"][ok])))
        out["cpu_baseline"] = {"value": round(n * nf / tc, 3), "unit": "columns·channels/s",
                               "cores": threads, "nproc": os.cpu_count(),
                               "cores_available": avail, "kind": "port",
                               "sample": f"first {n} of the {nx} columns x {nf} channels, "
                                         f"{tc:.1f} s"}
        out["max_abs_dtb_vs_oracle_K"] = dmax
        log(f"cpu baseline {n * nf / tc:.2f} col*ch/s on {threads} threads; "
            f"max |dTB| GPU vs oracle on the sample {dmax:.2e} K")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
