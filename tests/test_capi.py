"""The C-ABI boundary without a GPU: the library loads, exports every entry
point include/pamtra_amd.h declares, and the ctypes mirror matches the C
struct layouts."""

import ctypes as C
import os
import re
import subprocess

import pytest

from pamtra_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pamtra_amd.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pam_\w+)\s*\(", text, re.M)))


def test_header_declares_the_exported_set():
    assert _declared_functions() == sorted(_capi.EXPORTED_SYMBOLS)


@pytest.mark.parametrize("nummu", sorted(_capi.LIBRARIES))
def test_library_loads_and_exports_every_symbol(nummu):
    path = _capi.library_path(nummu)
    assert os.path.exists(path), "build the library first"
    lib = C.CDLL(path)
    for name in _declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", path],
                         capture_output=True, text=True, check=True).stdout
    for name in _declared_functions():
        assert re.search(r"\sT\s" + name + r"$", out, re.M), name


@pytest.mark.parametrize("nummu", sorted(_capi.LIBRARIES))
def test_version_and_device_count_without_gpu(nummu):
    lib = _capi.declare_library(C.CDLL(_capi.library_path(nummu)))
    assert lib.pam_version().startswith(b"pamtra_amd")
    assert lib.pam_device_count() >= 0
    assert lib.pam_nummu() == nummu


def test_angle_count_selects_the_library_and_result_shapes():
    from pamtra_amd import workloads
    pam, freqs = workloads.config5(ncol=2)
    m = _capi.marshal(pam.nmlSet, dict(pam.set, _nangles=pam._nangles), pam.df.data,
                      pam.df.data4D, pam.p)
    hr = _capi.alloc_results(m, len(freqs))
    assert m.nummu == 32 and hr.arrays["tb"].shape == (2, 1, 2, 64, 12, 2)
    assert hr.arrays["emissivity"].shape == (2, 1, 2, 12, 32)
    with pytest.raises(NotImplementedError):
        _capi.marshal(pam.nmlSet, dict(pam.set, _nangles=24), pam.df.data, pam.df.data4D, pam.p)


STRUCTS = {
    "pam_settings": _capi.PamSettings,
    "pam_hydro_desc": _capi.PamHydroDesc,
    "pam_hydro_4d": _capi.PamHydro4D,
    "pam_profile": _capi.PamProfile,
    "pam_results": _capi.PamResults,
}


def test_struct_layouts_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(){']
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"


def test_marshal_rejects_unsupported_options():
    from _scenario import build_pamtra
    pam = build_pamtra()
    nml = dict(pam.nmlSet, gas_mod="L93")
    with pytest.raises(NotImplementedError):
        _capi.marshal(nml, pam.set, pam.df.data, pam.df.data4D, pam.p)
    pam2 = build_pamtra()
    pam2.df.data["scat_name"][0] = "tmatrix"
    with pytest.raises(NotImplementedError):
        _capi.marshal(pam2.nmlSet, pam2.set, pam2.df.data, pam2.df.data4D, pam2.p)


def test_marshal_refuses_tensors_the_kernels_cannot_read():
    """A torch input must be float64 (int32 for counts), contiguous, of the
    expected element count and on the device; each fault is named."""
    import numpy as np
    import torch
    from _scenario import build_pamtra
    cases = [
        (lambda q: torch.tensor(q, dtype=torch.float32), "float64 is required"),
        (lambda q: torch.tensor(np.repeat(q, 2, axis=-1))[..., ::2], "not contiguous"),
        (lambda q: torch.tensor(q.ravel()[:-1].copy()), "elements"),
        (lambda q: torch.tensor(q), "device tensor"),
    ]
    for make, word in cases:
        pam = build_pamtra()
        q = np.ascontiguousarray(pam.p["hydro_q"])
        p = dict(pam.p)
        p["hydro_q"] = make(q)
        with pytest.raises(ValueError, match=word):
            _capi.marshal(pam.nmlSet, pam.set, pam.df.data, pam.df.data4D, p)


def test_product_path_fails_loudly_without_a_gpu(monkeypatch):
    """No CPU fallback: without a visible device runPamtra raises."""
    from pamtra_amd import _lib
    lib = _lib.library()
    if lib.pam_device_count() > 0:
        pytest.skip("a GPU is visible")
    from _scenario import build_pamtra
    pam = build_pamtra()
    pam.nmlSet["active"] = False
    with pytest.raises(_lib.PamLibraryError):
        pam.runPamtra([35.5])
