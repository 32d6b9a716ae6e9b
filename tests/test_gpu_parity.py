"""Parity of the HIP path (through the C-ABI) with the CPU oracle and the
reference's golden vectors.  FP64 bar: |dTB| <= 1e-3 K; discrete decisions
(doubling counts, Legendre orders) identical."""

import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from pamtra_amd import _capi, workloads
from _scenario import FREQUENCIES, build_pamtra, passive_nml

pytestmark = pytest.mark.gpu

TB_TOL = 1e-3  # K, FP64 device path vs oracle


def _gpu(pam, freqs, diagnostics=True):
    from pamtra_amd import _lib
    lib = _lib.library()
    ctx = _lib.get_context(0)
    m = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f), diagnostics=diagnostics)
    print(f"pam_run ncol={m.nx * m.ny} L={m.max_nlyrs} nf={len(f)}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    out = dict(hr.arrays)
    out["msg"] = _capi.results_message(hr)
    return out, st


def _cpu(pam, freqs):
    out, st, _ = oracle.run(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p, freqs,
                            nthreads=8, diagnostics=True)
    return out, st


def _compare(g, c, tol=TB_TOL):
    ok = np.isfinite(c["tb"])
    assert np.array_equal(ok, np.isfinite(g["tb"]))
    d = np.abs(g["tb"][ok] - c["tb"][ok])
    print(f"max |dTB| = {d.max():.3e} K", flush=True)
    assert d.max() <= tol
    np.testing.assert_array_equal(g["ndoubles"], c["ndoubles"])
    np.testing.assert_array_equal(g["nlegen"], c["nlegen"])
    m = np.isfinite(c["kextatmo"])
    np.testing.assert_allclose(g["kextatmo"][m], c["kextatmo"][m], rtol=1e-12)
    m = np.isfinite(c["kexttot"])
    np.testing.assert_allclose(g["kexttot"][m], c["kexttot"][m], rtol=1e-9, atol=1e-300)


def test_golden_through_pypamtra():
    """The reference's golden scenario through runPamtra on the GPU."""
    pam = build_pamtra()
    pam.nmlSet["active"] = False
    pam.runPamtra(FREQUENCIES)
    ref = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                             "rain_snow_reference.npz"))["tb"]
    np.testing.assert_allclose(pam.r["tb"], ref, rtol=1e-5)
    assert pam.r["tb"].shape == (1, 1, 2, 32, 2, 2)
    assert pam.r["Ze"] == -9999.
    np.testing.assert_allclose(pam.r["angles_deg"][:16], 180 - pam.r["angles_deg"][16:][::-1])


def test_golden_vs_oracle_with_diagnostics():
    pam = build_pamtra()
    g, st = _gpu(pam, FREQUENCIES)
    c, st2 = _cpu(pam, FREQUENCIES)
    assert st == 0 and st2 == 0
    _compare(g, c)


def test_config1_columns_vs_oracle():
    """The whole small configuration: 64 columns."""
    pam, freqs = workloads.config1(ncol=64, seed=1001)
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    _compare(g, c)


def test_config2_columns_vs_oracle():
    """Uneven layer counts, 7 channels, ocean (TESSEM2) and Lambertian land:
    the first 256 columns of the batched configuration."""
    pam, freqs = workloads.config2(ncol=256, seed=1002)
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    np.testing.assert_allclose(g["emissivity"], c["emissivity"], rtol=1e-12, atol=1e-15)
    _compare(g, c)


def test_adaptive_grid_and_soft_spheres_vs_oracle():
    """The golden's adaptive grid and exp_field_t soft spheres at more
    channels and contents."""
    pam = build_pamtra()
    pam.p["hydro_q"][0, 0, 0, 0] = 3e-4
    pam.p["hydro_q"][0, 0, 1, 1] = 2e-4
    freqs = [10.65, 23.8, 36.5, 89.0, 150.0, 183.31]
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0
    _compare(g, c)


def test_gauss_jordan_and_series_inverses_agree(monkeypatch):
    """Every (I - X)^-1 of the doubling and adding is a product-form series
    on the matrix cores, with Gauss-Jordan as its fallback.  Forcing
    Gauss-Jordan everywhere (PAM_INVERSE=gauss-jordan) must give the same
    TB to rounding (the FP64 noise floor of the whole chain, doubling from
    1e-6 optical depth, is ~1e-7 K), and both must match the oracle."""
    pam, freqs = workloads.config2(ncol=16, seed=21)
    a, st = _gpu(pam, freqs)
    monkeypatch.setenv("PAM_INVERSE", "gauss-jordan")
    b, st2 = _gpu(pam, freqs)
    monkeypatch.delenv("PAM_INVERSE")
    assert st == 0 and st2 == 0
    d = np.max(np.abs(a["tb"] - b["tb"]))
    print(f"series vs Gauss-Jordan max |dTB| = {d:.3e} K", flush=True)
    assert d <= 1e-6
    c, _ = _cpu(pam, freqs)
    _compare(b, c)
    # every unit through the queue of the Gauss-Jordan pass (the path of a
    # unit with an inverse whose series does not converge): the pass tries
    # the same series per inverse, with G formed explicitly where the hot
    # kernel applies the factors to T, so it agrees to the rounding of that
    # association
    monkeypatch.setenv("PAM_INVERSE", "queue-all")
    q, st3 = _gpu(pam, freqs)
    monkeypatch.delenv("PAM_INVERSE")
    assert st3 == 0
    dq = np.max(np.abs(q["tb"] - a["tb"]))
    print(f"queue pass vs hot kernel max |dTB| = {dq:.3e} K", flush=True)
    assert dq <= 1e-6
    _compare(q, c)


def test_optically_thick_scattering_vs_oracle():
    """Heavy rain and cloud (20x the contents) at 89-183 GHz: thick,
    strongly scattering layers, where the inverse series needs the most
    factors and the stack reflects most."""
    pam, freqs = workloads.config1(ncol=8, seed=33)
    pam.p["hydro_q"] *= 20.0
    freqs = [89.0, 150.0, 183.31]
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    _compare(g, c)


def test_bad_column_is_isolated():
    """A NaN temperature fails its own column with PAM_FATAL, the others are
    computed as if alone (status per column, no abort)."""
    pam, freqs = workloads.config1(ncol=6, seed=7)
    pam.p["temp_lev"][2, 0, :] = np.nan
    g, st = _gpu(pam, freqs)
    assert st == _capi.PAM_FATAL
    assert "nan" in g["msg"]
    assert (g["status"][2] == _capi.PAM_FATAL).all()
    good = [0, 1, 3, 4, 5]
    assert (g["status"][good] == 0).all()
    assert np.isnan(g["tb"][2]).all()
    pam2, _ = workloads.config1(ncol=6, seed=7)
    g2, _ = _gpu(pam2, freqs)
    np.testing.assert_array_equal(g["tb"][good], g2["tb"][good])


def test_column_chunking_is_bitwise_neutral(monkeypatch):
    """A workspace budget of a few layers per chunk (many launches per stage)
    gives the single-chunk result bit for bit."""
    pam, freqs = workloads.config2(ncol=12, seed=17)
    a, st = _gpu(pam, freqs)
    monkeypatch.setenv("PAM_WORKSPACE_GB", "0.004")  # ~25 scattering layers per chunk
    b, st2 = _gpu(pam, freqs)
    monkeypatch.delenv("PAM_WORKSPACE_GB")
    assert st == 0 and st2 == 0
    np.testing.assert_array_equal(a["tb"], b["tb"])
    np.testing.assert_array_equal(a["ndoubles"], b["ndoubles"])


def test_repeat_runs_are_bitwise_identical():
    pam, freqs = workloads.config2(ncol=8, seed=5)
    a, _ = _gpu(pam, freqs)
    b, _ = _gpu(pam, freqs)
    np.testing.assert_array_equal(a["tb"], b["tb"])


def test_device_entry_point_matches_host_entry_point():
    """pam_run_device on HBM-resident torch tensors gives pam_run's result."""
    torch = pytest.importorskip("torch")
    from pamtra_amd import _lib
    pam, freqs = workloads.config2(ncol=16, seed=11)
    host, _ = _gpu(pam, freqs, diagnostics=False)
    lib = _lib.library()
    ctx = _lib.get_context(0)
    prof = {}
    for k, v in pam.p.items():
        if isinstance(v, np.ndarray) and v.dtype.kind in "fiu" and k in (
                _capi.PROFILE_DOUBLE_KEYS + ["nlyrs", "sfc_type", "sfc_model", "obs_height"]):
            dt = torch.int32 if k in ("nlyrs", "sfc_type", "sfc_model") else torch.float64
            prof[k] = torch.as_tensor(np.ascontiguousarray(v), dtype=dt, device="cuda")
        else:
            prof[k] = v
    refl = np.frombuffer("".join(pam.p["sfc_refl"].reshape(-1)).encode(), dtype=np.uint8)
    prof["sfc_refl"] = torch.as_tensor(refl.copy(), device="cuda")
    m = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, prof)
    nx, ny, nout, nf = m.nx, m.ny, m.noutlevels, len(freqs)
    tb = torch.zeros((nx, ny, nout, 32, nf, 2), dtype=torch.float64, device="cuda")
    status = torch.zeros((nx, ny, nf), dtype=torch.int32, device="cuda")
    r = _capi.PamResults()
    r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
    r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))
    msg = C.create_string_buffer(256)
    r.msg = C.cast(msg, C.c_char_p)
    r.msglen = 256
    f, fp = _capi.freqs_array(freqs)
    st = lib.pam_run_device(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                            C.byref(m.profile), fp, nf, C.byref(r), None)
    assert st == 0, msg.value
    assert np.all(np.isfinite(host["tb"]))
    np.testing.assert_array_equal(tb.cpu().numpy(), host["tb"])


def test_large_particles_high_frequency():
    """Four categories up to 13 mm at 183 and 325 GHz: series of up to ~60
    terms, whose angle tables no longer fit the Mie kernel's LDS stage and are
    read from memory, and the longest Legendre series."""
    pam, _ = workloads.config3(ncol=6, seed=91)
    freqs = [183.31, 325.0]
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == st2, g["msg"]
    print("max nlegen", int(np.max(c["nlegen"])), flush=True)
    assert np.max(c["nlegen"]) > 80
    _compare(g, c)


def _levels(pam, heights):
    nx, ny = pam.p["ngridx"], pam.p["ngridy"]
    pam.p["noutlevels"] = len(heights[0])
    pam.p["obs_height"] = np.asarray(heights, dtype=float).reshape(nx, ny, len(heights[0]))
    return pam


@pytest.mark.parametrize("mode", [None, "gauss-jordan", "queue-all"])
def test_interior_output_levels_vs_oracle(monkeypatch, mode):
    """Output levels inside the column (radtran4.f90:375-422): those units
    go through k_adding_levels, which folds the stacks above and below each
    level.  Device vs oracle, and the top and surface levels of those units
    against the standard path's."""
    pam, freqs = workloads.config2(ncol=20, seed=91)
    hl = pam.p["hgt_lev"][:, 0, :]
    nl = pam.p["nlyrs"][:, 0]
    assert pam.p["obs_height"].shape[-1] == 2
    ref, st0 = _gpu(pam, freqs)
    assert st0 == 0
    heights = [[833000.0, 0.0, hl[c, nl[c] // 2], 0.5 * (hl[c, 2] + hl[c, 3]) + 1.0, 1500.0]
               for c in range(20)]
    pam = _levels(pam, heights)
    if mode:
        monkeypatch.setenv("PAM_INVERSE", mode)
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    _compare(g, c)
    d = np.max(np.abs(g["tb"][..., :2, :, :, :] - ref["tb"]))
    print(f"top/surface levels through the level pass vs the standard path: {d:.2e} K", flush=True)
    assert d <= 1e-7


def test_level_asked_twice_and_mixed_precision():
    """Two observation heights above the column (both the top level) and a
    mixed-precision run with an interior level."""
    pam, freqs = workloads.config2(ncol=6, seed=92)
    ref, _ = _gpu(pam, freqs)
    pam = _levels(pam, [[833000.0, 0.0, 900000.0]] * 6)
    g, st = _gpu(pam, freqs)
    assert st == 0
    np.testing.assert_allclose(g["tb"][..., 2, :, :, :], ref["tb"][..., 0, :, :, :], rtol=0,
                               atol=1e-7)
    np.testing.assert_allclose(g["tb"][..., :2, :, :, :], ref["tb"], rtol=0, atol=1e-7)
    hl = pam.p["hgt_lev"][:, 0, :]
    pam = _levels(pam, [[833000.0, 0.0, hl[c, 7]] for c in range(6)])
    a, _ = _gpu(pam, freqs)
    pam.set["precision"] = "mixed"
    b, st3 = _gpu(pam, freqs)
    pam.set["precision"] = "fp64"
    assert st3 == 0
    assert np.max(np.abs(a["tb"] - b["tb"])) <= 0.05


class _PtrAttr(C.Structure):
    _fields_ = [("type", C.c_int), ("device", C.c_int), ("devicePointer", C.c_void_p),
                ("hostPointer", C.c_void_p), ("isManaged", C.c_int), ("allocationFlags", C.c_uint)]


def _device_profile(pam, torch):
    prof = {}
    for k, v in pam.p.items():
        if isinstance(v, np.ndarray) and v.dtype.kind in "fiu" and k in (
                _capi.PROFILE_DOUBLE_KEYS + ["nlyrs", "sfc_type", "sfc_model", "obs_height"]):
            dt = torch.int32 if k in ("nlyrs", "sfc_type", "sfc_model") else torch.float64
            prof[k] = torch.as_tensor(np.ascontiguousarray(v), dtype=dt, device="cuda")
        else:
            prof[k] = v
    refl = np.frombuffer("".join(pam.p["sfc_refl"].reshape(-1)).encode(), dtype=np.uint8)
    prof["sfc_refl"] = torch.as_tensor(refl.copy(), device="cuda")
    return prof


def _run_device(lib, ctx, m, freqs, torch):
    nx, ny, nout, nf = m.nx, m.ny, m.noutlevels, len(freqs)
    tb = torch.zeros((nx, ny, nout, 32, nf, 2), dtype=torch.float64, device="cuda")
    status = torch.zeros((nx, ny, nf), dtype=torch.int32, device="cuda")
    r = _capi.PamResults()
    r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
    r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))
    msg = C.create_string_buffer(256)
    r.msg = C.cast(msg, C.c_char_p)
    r.msglen = 256
    f, fp = _capi.freqs_array(freqs)
    st = lib.pam_run_device(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d),
                            C.byref(m.profile), fp, nf, C.byref(r), None)
    return st, tb, msg.value


def test_device_entry_with_4d_descriptors_and_host_array_refusal():
    """The 4D descriptor (per-column snow density) read from HBM gives
    pam_run's result; the same call with that array left in host memory is
    refused before any launch.  The refusal is exercised only after the HIP
    runtime has been seen to classify the host array as host memory."""
    torch = pytest.importorskip("torch")
    from pamtra_amd import _lib
    pam, freqs = workloads.config3(ncol=6, seed=13)
    freqs = freqs[:3]
    assert "rho_ms" in pam.df.data4D
    host, _ = _gpu(pam, freqs, diagnostics=False)
    lib = _lib.library()
    ctx = _lib.get_context(0)
    prof = _device_profile(pam, torch)
    d4 = {k: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device="cuda")
          for k, v in pam.df.data4D.items()}
    m = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, d4, prof)
    print("pam_run_device with HBM 4D descriptors", flush=True)
    st, tb, msg = _run_device(lib, ctx, m, freqs, torch)
    assert st == 0, msg
    assert np.all(np.isfinite(host["tb"]))
    np.testing.assert_array_equal(tb.cpu().numpy(), host["tb"])
    h4 = np.ascontiguousarray(pam.df.data4D["rho_ms"], dtype=np.float64)
    hip = C.CDLL("libamdhip64.so.7")
    a = _PtrAttr()
    e = hip.hipPointerGetAttributes(C.byref(a), C.c_void_p(h4.ctypes.data))
    hip.hipGetLastError()
    if e == 0 and (a.type in (2, 3) or (a.type == 1 and a.devicePointer)):
        pytest.skip("the HIP runtime reports this host array as device-accessible")
    mh = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, {"rho_ms": h4}, prof)
    print("pam_run_device with a host 4D array (refused before launch)", flush=True)
    st2, _, msg2 = _run_device(lib, ctx, mh, freqs, torch)
    assert st2 == _capi.PAM_FATAL
    assert b"hydro4d.rho_ms is not device memory" in msg2


def test_add_obs_height_to_layer_through_pypamtra():
    """nmlSet add_obs_height_to_layer (core.py:1553, :1908-1951): runPamtra
    inserts the observation height as a level and outputs there; the device
    matches the oracle on the regridded profile."""
    pam, freqs = workloads.config1(ncol=4, seed=98)
    nx, ny = pam.p["ngridx"], pam.p["ngridy"]
    pam.p["noutlevels"] = 3
    pam._shape3Dout = (nx, ny, 3)
    pam.p["obs_height"] = np.array([[833000.0, 2150.0, 0.0]] * nx).reshape(nx, ny, 3)
    pam.nmlSet["active"] = False
    pam.nmlSet["add_obs_height_to_layer"] = True
    pam.runPamtra(freqs)
    assert pam.p["max_nlyrs"] == 31 and 2150.0 in pam.p["hgt_lev"][0, 0]
    pam.nmlSet["add_obs_height_to_layer"] = False
    c, st = _cpu(pam, freqs)
    assert st == 0
    d = np.max(np.abs(pam.r["tb"] - c["tb"]))
    print(f"max |dTB| = {d:.3e} K", flush=True)
    assert d <= TB_TOL


def test_fastem_ocean_vs_oracle():
    """sfc_model = 1 (FASTEM, sfc_optics.f90:57-58) on ocean columns with
    wind: device emissivities and TB against the oracle's."""
    pam, freqs = workloads.config2(ncol=16, seed=99)
    ocean = pam.p["sfc_type"] == 0
    assert ocean.any()
    pam.p["sfc_model"][ocean] = 1
    pam.p["wind10u"][:] = 7.0
    pam.p["wind10v"][:] = -3.0
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == 0 and st2 == 0, (g["msg"], c["msg"])
    np.testing.assert_allclose(g["emissivity"], c["emissivity"], rtol=1e-12, atol=1e-15)
    _compare(g, c)


def test_iq_output_vs_oracle():
    """outpol = 'IQ' (rt_utilities.f90:687-737): I and Q instead of V and H."""
    pam, freqs = workloads.config2(ncol=16, seed=1010)
    pam.nmlSet["outpol"] = "IQ"
    g, st = _gpu(pam, freqs)
    c, st2 = _cpu(pam, freqs)
    assert st == st2 and st in (0, 1), (g["msg"], c["msg"])
    _compare(g, c)
    pam.nmlSet["outpol"] = "VH"
    v, _ = _gpu(pam, freqs)
    # the same radiances: B(V) + B(H) = 2 B(I) and B(V) - B(H) = 2 B(Q), with
    # B the Planck radiance of CONVERT_OUTPUT (sign kept for negative Q)
    wl = 299792458.0 / (np.asarray(freqs) * 1e3)
    wl = wl[None, None, None, None, :]

    def planck(t):
        a = np.abs(t)
        return np.sign(t) * 1.1911e8 / wl ** 5 / (np.exp(1.4388e4 / (wl * a)) - 1.0)

    bi, bq = planck(g["tb"][..., 0]), planck(g["tb"][..., 1])
    bv, bh = planck(v["tb"][..., 0]), planck(v["tb"][..., 1])
    np.testing.assert_allclose(bv + bh, 2.0 * bi, rtol=1e-9)
    np.testing.assert_allclose(bv - bh, 2.0 * bq, rtol=1e-6, atol=1e-9 * np.max(np.abs(bi)))
