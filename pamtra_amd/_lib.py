"""Loads libpamtra_amd.so (built in-tree) and holds one context per device.

Fails loudly when the library is missing: there is no CPU path behind it.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

from . import _capi

_lib = None
_ctx = {}
_lock = threading.Lock()


class PamLibraryError(RuntimeError):
    pass


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch bundles its own libamdhip64
        # (soname libamdhip64.so.7) and loads it by a different file name, so
        # it must be loaded first for this library's dependency to resolve to
        # the same copy.  Importing torch does not initialise the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        path = _capi.LIBRARY_PATH
        if not os.path.exists(path):
            raise PamLibraryError(
                f"{path} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        _lib = _capi.declare_library(C.CDLL(path))
    return _lib


def get_context(device: int = 0) -> C.c_void_p:
    """Context bound to `device`, created on first use."""
    with _lock:
        if device not in _ctx:
            lib = library()
            if lib.pam_device_count() <= device:
                raise PamLibraryError(f"no HIP device {device} visible to libpamtra_amd")
            h = C.c_void_p()
            st = lib.pam_context_create(C.byref(h), device)
            if st != 0 or not h.value:
                raise PamLibraryError(f"pam_context_create failed on device {device}")
            _ctx[device] = h
        return _ctx[device]
