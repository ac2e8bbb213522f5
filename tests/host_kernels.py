"""Builds and binds tests/native/host_kernels.cpp (test infrastructure)."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "native" / "host_kernels.cpp"
OUT = HERE / "_build" / "libhost_kernels.so"
HDR = HERE.parent / "crushrs_amd" / "csrc" / "device_common.hpp"

_lib = None
f32p = C.POINTER(C.c_float)


def load():
    global _lib
    if _lib is None:
        OUT.parent.mkdir(exist_ok=True)
        stale = not OUT.exists() or OUT.stat().st_mtime < max(SRC.stat().st_mtime, HDR.stat().st_mtime)
        if stale:
            hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
            subprocess.run([hipcc, "-O2", "-x", "hip", "--offload-arch=gfx950", "-shared", "-fPIC",
                            "-o", str(OUT), str(SRC)], check=True, capture_output=True)
        lib = C.CDLL(str(OUT))
        lib.hk_hencky.argtypes = [f32p, C.c_float, C.c_float, f32p]
        lib.hk_honeycomb.argtypes = [f32p, C.c_float, f32p, C.c_float, f32p, f32p]
        lib.hk_eigen.argtypes = [f32p, f32p, f32p]
        _lib = lib
    return _lib


def fp(a):
    return a.ctypes.data_as(f32p)
