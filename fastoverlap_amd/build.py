"""Build libfastoverlap_hip.so in-tree for gfx950.

Each source is compiled to an object with hipcc (in parallel) and linked into
``fastoverlap_amd/libfastoverlap_hip.so``.  Usage::

    python -m fastoverlap_amd.build          # or: python fastoverlap_amd/build.py
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfastoverlap_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FO_OFFLOAD_ARCH", "gfx950")

SOURCES = ["fo_spherical.hip", "fo_periodic.hip", "fo_api.cpp", "fo_server.cpp", "fo_f90.cpp",
           "fo_noalloc.hip", "fo_peaks.hip", "fo_harmonic.hip"]
FORTRAN = os.path.join(HERE, "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay disabled (see the module docs)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    if os.path.exists(FLANG):
        build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = False) -> str:
    """Compile the iso_c_binding modules and the Fortran test caller
    (fortran/_build/test_align, linked against the in-tree library)."""
    out = os.path.join(FORTRAN, "_build")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(out, "test_align")
    cmd = [FLANG, "-O2", "-J", out, os.path.join(FORTRAN, "fastoverlap_hip.f90"),
           os.path.join(FORTRAN, "test_align.f90"), "-o", exe, "-L", HERE,
           "-lfastoverlap_hip", f"-Wl,-rpath,$ORIGIN/../..", f"-Wl,-rpath,{HERE}"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=out)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    try:
        build(verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
