"""ctypes binding of the f64 CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the oracle is the checker the GPU path is
compared against.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from crushrs_amd import _abi

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return ORACLE_LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            build()
        lib = C.CDLL(str(ORACLE_LIB))
        f64p, i32p, u8p = _abi.f64p, _abi.i32p, _abi.u8p
        lib.oracle_hex_geometry.argtypes = [f64p, C.c_double, C.c_double, C.c_double, C.c_double, f64p, f64p,
                                            f64p, f64p, f64p]
        lib.oracle_dense_hourglass_k.argtypes = [f64p, C.c_double, C.c_double, C.c_double, f64p]
        lib.oracle_honeycomb_params.argtypes = [C.POINTER(_abi.Material), f64p, f64p]
        lib.oracle_knot_yield_at.argtypes = [f64p, C.c_double, f64p, f64p]
        lib.oracle_knot_yield_at.restype = None
        lib.oracle_plasticity_yield_at.argtypes = [C.POINTER(_abi.Material), C.c_double, f64p, f64p]
        lib.oracle_plasticity_yield_at.restype = None
        lib.oracle_honeycomb_update.argtypes = [C.c_double, f64p, f64p, f64p, f64p, f64p]
        lib.oracle_honeycomb_update.restype = None
        lib.oracle_eigen3.argtypes = [f64p, f64p, f64p]
        lib.oracle_eigen3.restype = None
        lib.oracle_hencky_elastic.argtypes = [C.c_double, C.c_double, f64p, f64p]
        lib.oracle_hencky_elastic.restype = None
        lib.oracle_char_length.argtypes = [f64p, f64p, f64p]
        lib.oracle_char_length.restype = None
        lib.oracle_element_force.argtypes = [f64p, f64p, f64p, C.c_double, f64p, f64p, C.c_int, C.c_double,
                                             C.c_double, f64p, f64p, f64p, f64p]
        lib.oracle_run.argtypes = [C.POINTER(_abi.ModelView), C.POINTER(_abi.Results)]
        lib.oracle_run_steps.argtypes = [C.POINTER(_abi.ModelView), C.c_int32, f64p, f64p, i32p, u8p, f64p, i32p]
        lib.oracle_contact_apply.argtypes = [C.POINTER(_abi.ModelView), f64p, f64p, C.c_double, f64p, i32p]
        lib.oracle_run_lanes.argtypes = [C.POINTER(_abi.ModelView), C.c_int32, C.POINTER(_abi.Results)]
        lib.oracle_run_steps_lanes.argtypes = [C.POINTER(_abi.ModelView), C.c_int32, C.c_int32, f64p, f64p, i32p,
                                               u8p, f64p, i32p]
        lib.oracle_run_displacements.argtypes = [C.POINTER(_abi.ModelView), C.c_int32, C.c_int32, f64p, i32p]
        lib.oracle_cfc_filter.argtypes = [f64p, C.c_int32, C.c_double, C.c_double, f64p]
        lib.oracle_cfc_filter.restype = None
        lib.oracle_resample.argtypes = [f64p, f64p, C.c_int32, C.c_double, C.c_double, C.c_double, f64p, C.c_int32]
        lib.oracle_integrate.argtypes = [f64p, f64p, C.c_int32, C.c_double, f64p]
        lib.oracle_integrate.restype = None
        lib.oracle_set_element_hook.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        lib.oracle_set_element_hook.restype = None
        lib.lane_blocks_create.argtypes = [C.c_int]
        lib.lane_blocks_create.restype = C.c_void_p
        lib.lane_blocks_free.argtypes = [C.c_void_p]
        lib.lane_blocks_free.restype = None
        lib.lane_blocks_set.argtypes = [C.c_void_p, C.c_int, i32p, f64p, C.c_double, f64p, f64p, C.c_double, f64p,
                                        f64p, f64p, C.c_int]
        lib.lane_blocks_set.restype = None
        lib.lane_blocks_forces.argtypes = [C.c_void_p, f64p, f64p, f64p, C.c_int]
        lib.lane_blocks_eroded.argtypes = [C.c_void_p, C.c_int]
        lib.lane_blocks_state.argtypes = [C.c_void_p, C.c_int, f64p]
        lib.lane_blocks_state.restype = None
        lib.lane_blocks_out.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        lib.lane_blocks_out.restype = None
        _lib = lib
    return _lib


def ptr(a: np.ndarray):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_abi.f64p)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_abi.i32p)
    if a.dtype == np.uint8:
        return a.ctypes.data_as(_abi.u8p)
    raise TypeError(a.dtype)


def run(model):
    """oracle_run on a crushrs_amd.Model -> crushrs_amd.Results."""
    import crushrs_amd as ca
    r, u, v, m, g = ca.alloc_model_results(model)
    rc = load().oracle_run(model.ptr, C.byref(r))
    assert rc == 0, f"oracle_run failed: {rc}"
    return ca.to_results(r, u, v, m, g, model.settings.pulse_dt)


def run_lanes(model, threads: int = 1):
    """oracle_run_lanes (the reference's default f32 AVX2 element path)."""
    import crushrs_amd as ca
    r, u, v, m, g = ca.alloc_model_results(model)
    rc = load().oracle_run_lanes(model.ptr, threads, C.byref(r))
    assert rc == 0, f"oracle_run_lanes failed: {rc}"
    return ca.to_results(r, u, v, m, g, model.settings.pulse_dt)


def run_steps(model, n_steps: int, n_contact_nodes: int):
    n3 = 3 * model.n_nodes
    fi = np.zeros((n_steps, n3))
    fx = np.zeros((n_steps, n3))
    hit = np.zeros((n_steps, max(n_contact_nodes, 1)), np.int32)
    er = np.zeros((n_steps, model.n_hexes), np.uint8)
    dt = np.zeros(n_steps)
    done = C.c_int32()
    rc = load().oracle_run_steps(model.ptr, n_steps, ptr(fi), ptr(fx), ptr(hit), ptr(er), ptr(dt), C.byref(done))
    assert rc == 0
    s = done.value
    return fi[:s], fx[:s], hit[:s, :n_contact_nodes], er[:s], dt[:s]
