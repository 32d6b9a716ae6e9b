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


def test_library_loads_and_exports_every_symbol():
    assert os.path.exists(_capi.LIBRARY_PATH), "build the library first"
    lib = C.CDLL(_capi.LIBRARY_PATH)
    for name in _declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIBRARY_PATH],
                         capture_output=True, text=True, check=True).stdout
    for name in _declared_functions():
        assert re.search(r"\sT\s" + name + r"$", out, re.M), name


def test_version_and_device_count_without_gpu():
    lib = _capi.declare_library(C.CDLL(_capi.LIBRARY_PATH))
    assert lib.pam_version().startswith(b"pamtra_amd")
    assert lib.pam_device_count() >= 0


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
