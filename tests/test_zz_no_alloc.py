"""pam_run_device_noalloc: the allocation-free variant of pam_run_device,
whose workspace is the library's fixed per-device scratch array.  Its results
must equal pam_run_device's bit for bit (same kernels; the column chunking is
bitwise neutral), on two calls in a row, the second smaller than the first
(the scratch keeps the first call's contents), and a call that cannot fit
the scratch is refused with a message."""

import ctypes as C

import numpy as np
import pytest

from pamtra_amd import _capi, workloads
from _scenario import passive_nml

pytestmark = pytest.mark.gpu


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


def _call(entry, m, freqs, torch):
    nx, ny, nout, nf = m.nx, m.ny, m.noutlevels, len(freqs)
    tb = torch.full((nx, ny, nout, 32, nf, 2), -1.0, dtype=torch.float64, device="cuda")
    status = torch.full((nx, ny, nf), -1, dtype=torch.int32, device="cuda")
    r = _capi.PamResults()
    r.tb = C.cast(C.c_void_p(tb.data_ptr()), C.POINTER(C.c_double))
    r.status = C.cast(C.c_void_p(status.data_ptr()), C.POINTER(C.c_int32))
    msg = C.create_string_buffer(256)
    r.msg = C.cast(msg, C.c_char_p)
    r.msglen = 256
    f, fp = _capi.freqs_array(freqs)
    st = entry(C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile), fp, nf,
               C.byref(r), None)
    return st, tb.cpu().numpy(), status.cpu().numpy(), msg.value


def _both(lib, ctx, pam, freqs, torch):
    m = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D,
                      _device_profile(pam, torch))
    ncol = m.nx * m.ny
    print(f"pam_run_device ncol={ncol} nf={len(freqs)}", flush=True)
    ref = _call(lambda *a: lib.pam_run_device(ctx, *a), m, freqs, torch)
    print(f"pam_run_device_noalloc ncol={ncol} nf={len(freqs)}", flush=True)
    out = _call(lambda *a: lib.pam_run_device_noalloc(ctx, *a), m, freqs, torch)
    return ref, out


def test_noalloc_matches_pam_run_device_on_two_calls():
    torch = pytest.importorskip("torch")
    from pamtra_amd import _lib
    lib, ctx = _lib.library(), _lib.get_context(0)
    per_item = lib.pam_workspace_bytes(1024, 120, 2, 80, 7, 1, 2) - \
        lib.pam_workspace_bytes(1024, 120, 2, 80, 7, 1, 1)
    # the first call's optics exceed the scratch several times over: chunks
    pam, freqs = workloads.config2(ncol=1024, seed=21)
    assert 1024 * 10 * per_item > 2 * _capi.NOALLOC_SCRATCH_BYTES
    ref, out = _both(lib, ctx, pam, freqs, torch)
    # Lambertian land columns warn (sfc_optics.f90:79-95), as in the reference
    assert ref[0] in (_capi.PAM_SUCCESS, _capi.PAM_WARNING), ref[3]
    assert out[0] == ref[0] and out[3] == ref[3]
    assert np.all(np.isfinite(ref[1]))
    np.testing.assert_array_equal(out[1], ref[1])
    np.testing.assert_array_equal(out[2], ref[2])
    # a smaller second call reuses the scratch left by the first
    pam2, freqs2 = workloads.config2(ncol=24, seed=22)
    ref2, out2 = _both(lib, ctx, pam2, freqs2[:4], torch)
    assert ref2[0] in (_capi.PAM_SUCCESS, _capi.PAM_WARNING), ref2[3]
    assert out2[0] == ref2[0] and out2[3] == ref2[3]
    np.testing.assert_array_equal(out2[1], ref2[1])
    np.testing.assert_array_equal(out2[2], ref2[2])
    assert np.all(np.isfinite(out2[1]))


def test_noalloc_refuses_a_column_larger_than_the_scratch():
    """One column whose optics alone exceed the scratch: 107 cloudy layers
    at 256 channels (the refusal comes before any per-layer kernel)."""
    torch = pytest.importorskip("torch")
    from pamtra_amd import _lib
    lib, ctx = _lib.library(), _lib.get_context(0)
    pam, _ = workloads.config2(ncol=1, seed=26)
    L = int(pam.p["nlyrs"][0, 0])
    pam.p["hydro_q"][0, 0, :L, 0] = 1e-4
    freqs = list(np.linspace(10.0, 300.0, 256))
    per_item = lib.pam_workspace_bytes(1, L, 2, 80, 256, 1, 2) - \
        lib.pam_workspace_bytes(1, L, 2, 80, 256, 1, 1)
    assert L * per_item > _capi.NOALLOC_SCRATCH_BYTES
    m = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D,
                      _device_profile(pam, torch))
    print(f"pam_run_device_noalloc oversized: L={L} nf=256", flush=True)
    st, _, _, msg = _call(lambda *a: lib.pam_run_device_noalloc(ctx, *a), m, freqs, torch)
    assert st == _capi.PAM_FATAL
    assert b"does not fit" in msg, msg
