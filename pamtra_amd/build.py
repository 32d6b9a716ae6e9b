"""Builds the library in-tree with hipcc for gfx950, in two variants of the
same sources: libpamtra_amd.so (16 angles per hemisphere, the reference's
compile-time nummu) and libpamtra_amd_n64.so (PAM_NUMMU=32: 32 angles,
64x64 operators).

Usage: python -m pamtra_amd.build   (or __graft_entry__.build())
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpamtra_amd.so")
VARIANTS = {16: "libpamtra_amd.so", 32: "libpamtra_amd_n64.so"}
SOURCES = ["pam_column.hip", "pam_dsd.hip", "pam_mie.hip", "pam_phase.hip", "pam_rt.hip",
           "pam_tmatrix.hip", "pam_capi.hip"]
ARCH = os.environ.get("PAM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-lambda-capture",
    # packed FP32 instructions stay off: plain scalar FP32/FP64 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(job) -> str:
    src, nummu = job
    obj = os.path.join(BUILD, f"{os.path.splitext(src)[0]}_mu{nummu}.o")
    cmd = [hipcc(), *FLAGS, f"-DPAM_NUMMU={nummu}", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src} (PAM_NUMMU={nummu}):\n{r.stderr}")
    if r.stderr.strip():
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln
                 and "not a recognized feature" not in ln]
        if any(ln.strip() for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(verbose: bool = True) -> str:
    """Builds both variants; returns the path of the 16-angle library."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = [(src, mu) for mu in VARIANTS for src in SOURCES]
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, jobs))
    for k, (mu, name) in enumerate(VARIANTS.items()):
        lib = os.path.join(HERE, name)
        tmp = lib + ".tmp"
        part = objs[k * len(SOURCES):(k + 1) * len(SOURCES)]
        # -Bsymbolic: each variant binds its own symbols when both are loaded
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", *part,
               "-o", tmp]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr}")
        os.replace(tmp, lib)
        if verbose:
            print(f"built {lib}")
    return LIB


if __name__ == "__main__":
    build()
