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
        f64p = C.POINTER(C.c_double)
        lib.hk_f32_round.argtypes = [C.POINTER(C.c_double), C.c_int, f32p, f32p]
        lib.hk_f32_round.restype = None
        lib.hk_char_length.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.hk_char_length.restype = None
        ip = C.POINTER(C.c_int)
        lib.hk_face_check.argtypes = [f64p, C.c_int, C.c_double, f64p, C.c_int, f64p, ip, ip, ip]
        lib.hk_face_check.restype = None
        lib.hk_hook_create.argtypes = [C.c_int]
        lib.hk_hook_create.restype = C.c_void_p
        lib.hk_hook_free.argtypes = [C.c_void_p]
        lib.hk_hook_free.restype = None
        lib.hk_hook_set.argtypes = [C.c_void_p, C.c_int, f64p, C.c_double, f64p, f64p, C.c_float, C.c_float, f64p,
                                    C.c_float]
        lib.hk_hook_set.restype = None
        lib.hk_hook_enable.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.hk_hook_enable.restype = None
        lib.hk_hook64_create.argtypes = [C.c_void_p, C.c_int]
        lib.hk_hook64_create.restype = C.c_void_p
        lib.hk_hook64_free.argtypes = [C.c_void_p]
        lib.hk_hook64_free.restype = None
        _lib = lib
    return _lib


def fp(a):
    return a.ctypes.data_as(f32p)


def element_hook(model, parts_sel=None):
    """A context for oracle_set_element_hook that computes element forces
    with the product's f32 element arithmetic (hk_hook), set up from the
    oracle's own f64 geometry of `model`.  Returns (ctx, fn_pointer)."""
    import numpy as np
    import oracle_ffi as of
    from crushrs_amd import _abi
    L, O = load(), of.load()
    v = model.view
    E = v.n_hexes
    parts_on = set() if parts_sel is None else set(parts_sel)
    ctx = L.hk_hook_create(E)
    X = np.ctypeslib.as_array(v.nodes, shape=(v.n_nodes, 3))
    hexes = np.ctypeslib.as_array(v.hexes, shape=(E, 8))
    parts = np.ctypeslib.as_array(v.hex_part, shape=(E,))
    knots = {}
    for p in range(v.n_parts):
        k, cap = np.zeros(24), np.zeros(3)
        if v.materials[p].model == _abi.MAT_HONEYCOMB:
            assert O.oracle_honeycomb_params(C.byref(v.materials[p]), of.ptr(k), of.ptr(cap)) == 0
        knots[p] = (k, cap)
    grad, vol, lump, gam, H = np.zeros(24), np.zeros(1), np.zeros(8), np.zeros(32), np.zeros(144)
    hg = model.settings.hourglass
    for e in range(E):
        mt = v.materials[parts[e]]
        x = np.ascontiguousarray(X[hexes[e]].ravel())
        assert O.oracle_hex_geometry(of.ptr(x), mt.youngs_modulus, mt.poisson_ratio, mt.density, hg, of.ptr(grad),
                                     of.ptr(vol), of.ptr(lump), of.ptr(gam), of.ptr(H)) == 0
        k, cap = knots[parts[e]]
        L.hk_hook_set(ctx, e, of.ptr(grad), vol[0], of.ptr(gam), of.ptr(H), mt.youngs_modulus, mt.poisson_ratio,
                      of.ptr(k), cap[1] if cap[1] > 0 else 1.0)
        if parts_sel is not None:
            L.hk_hook_enable(ctx, e, int(parts[e]) in parts_on)
    return ctx, C.cast(L.hk_hook, C.c_void_p)
