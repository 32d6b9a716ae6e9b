"""Loads the library (built in-tree: libpamtra_amd.so for 16 angles per
hemisphere, libpamtra_amd_n64.so for 32) and holds one context per device.

Fails loudly when the library is missing: there is no CPU path behind it.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

from . import _capi

_libs = {}
_ctx = {}
_lock = threading.Lock()


class PamLibraryError(RuntimeError):
    pass


def library(nummu: int = _capi.NUMMU) -> C.CDLL:
    """The library built for `nummu` quadrature angles per hemisphere."""
    if nummu not in _libs:
        # One HIP runtime per process: torch bundles its own libamdhip64
        # (soname libamdhip64.so.7) and loads it by a different file name, so
        # it must be loaded first for this library's dependency to resolve to
        # the same copy.  Importing torch does not initialise the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        path = _capi.library_path(nummu)
        if not os.path.exists(path):
            raise PamLibraryError(
                f"{path} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = _capi.declare_library(C.CDLL(path))
        if lib.pam_nummu() != nummu:
            raise PamLibraryError(f"{path} is built for {lib.pam_nummu()} angles, not {nummu}")
        _libs[nummu] = lib
    return _libs[nummu]


def get_context(device: int = 0, nummu: int = _capi.NUMMU) -> C.c_void_p:
    """Context of the `nummu` library bound to `device`, created on first use."""
    with _lock:
        if (nummu, device) not in _ctx:
            lib = library(nummu)
            if lib.pam_device_count() <= device:
                raise PamLibraryError(f"no HIP device {device} visible to libpamtra_amd")
            h = C.c_void_p()
            st = lib.pam_context_create(C.byref(h), device)
            if st != 0 or not h.value:
                raise PamLibraryError(f"pam_context_create failed on device {device}")
            _ctx[(nummu, device)] = h
        return _ctx[(nummu, device)]
