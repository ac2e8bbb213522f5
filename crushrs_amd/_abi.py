"""ctypes mirror of include/crushrs.h and include/crushrs_scenario.h."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

MAX_KNOTS = 8
MAX_TABLE = 8

CRUSH_OK = 0
CRUSH_E_INPUT = 1
CRUSH_E_NONFINITE = 2
CRUSH_E_OOM = 3
CRUSH_E_DEVICE = 4
CRUSH_E_UNSUPPORTED = 5
CRUSH_E_BUSY = 6
CRUSH_E_NOTFOUND = 7

MAT_ELASTIC, MAT_HONEYCOMB, MAT_J2, MAT_FOAM = 0, 1, 2, 3

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


class Material(C.Structure):
    _fields_ = [
        ("youngs_modulus", C.c_double),
        ("poisson_ratio", C.c_double),
        ("density", C.c_double),
        ("model", C.c_int32),
        ("n_curve", C.c_int32),
        ("yield_stress", C.c_double),
        ("hardening", C.c_double),
        ("curve", (C.c_double * 2) * MAX_KNOTS),
        ("has_densification", C.c_int32),
        ("reserved0", C.c_int32),
        ("densification", C.c_double * 2),
        ("transverse_factor", C.c_double),
    ]


class Contact(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32),
        ("n_faces", C.c_int32),
        ("nodes", i32p),
        ("node_scale", f64p),
        ("faces", i32p),
        ("stiffness", C.c_double),
        ("max_distance", C.c_double),
        ("pair", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("end_time", C.c_double),
        ("dt_scale", C.c_double),
        ("time_step", C.c_double),
        ("hourglass", C.c_double),
        ("velocity_damping", C.c_double),
        ("mass_scaling", C.c_double),
        ("adaptive_check_steps", C.c_int32),
        ("separation_steps", C.c_int32),
        ("pulse_dt", C.c_double),
        ("history_steps", C.c_int32),
        ("frame_steps", C.c_int32),
    ]


class Group(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32),
        ("n_front", C.c_int32),
        ("nodes", i32p),
        ("front", i32p),
        ("accel_node", C.c_int32),
        ("axis", C.c_int32),
        ("frame_origin", C.c_int32),
        ("frame_x", C.c_int32),
        ("heading", C.c_int32),
        ("frame_plane", C.c_int32),
    ]


class ModelView(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32),
        ("n_hexes", C.c_int32),
        ("n_parts", C.c_int32),
        ("n_faces", C.c_int32),
        ("n_fixed", C.c_int32),
        ("n_contacts", C.c_int32),
        ("n_groups", C.c_int32),
        ("reserved0", C.c_int32),
        ("nodes", f64p),
        ("hexes", i32p),
        ("hex_part", i32p),
        ("materials", C.POINTER(Material)),
        ("initial_velocity", f64p),
        ("fixed_nodes", i32p),
        ("faces", i32p),
        ("contacts", C.POINTER(Contact)),
        ("groups", C.POINTER(Group)),
        ("settings", Settings),
    ]


class GroupResult(C.Structure):
    _fields_ = [
        ("mass", C.c_double),
        ("delta_v", C.c_double * 3),
        ("crush_end", C.c_double),
        ("crush_max", C.c_double),
        ("pulse_delta_v", C.c_double),
    ]


class Results(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("steps", C.c_int64),
        ("eroded_elements", C.c_int64),
        ("wall_ms", C.c_double),
        ("device_ms", C.c_double),
        ("added_mass_max", C.c_double),
        ("separated", C.c_int32),
        ("n_groups", C.c_int32),
        ("displacement", f64p),
        ("velocity", f64p),
        ("masses", f64p),
        ("groups", C.POINTER(GroupResult)),
        ("pulse", f64p),
        ("pulse_capacity", C.c_int32),
        ("pulse_samples", C.c_int32),
        ("history_time", f64p),
        ("history_step", C.POINTER(C.c_int64)),
        ("history", f64p),
        ("accel_history", f64p),
        ("history_capacity", C.c_int32),
        ("history_samples", C.c_int32),
        ("frame_time", f64p),
        ("frame_displacement", C.POINTER(C.c_float)),
        ("frame_plastic_strain", C.POINTER(C.c_float)),
        ("frame_stress", C.POINTER(C.c_float)),
        ("frame_eroded", u8p),
        ("frame_capacity", C.c_int32),
        ("frame_samples", C.c_int32),
    ]


PART_FIELDS = 11   # CRUSH_PART_FIELDS
ACCEL_FIELDS = 24  # CRUSH_ACCEL_FIELDS


class Vehicle(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("mass", C.c_double),
        ("size", C.c_double * 3),
        ("crush_zone_length", C.c_double),
        ("curve", C.c_double * 2),
        ("modulus", C.c_double),
        ("body_modulus", C.c_double),
        ("has_side_curve", C.c_int32),
        ("has_side_modulus", C.c_int32),
        ("side_curve", C.c_double * 2),
        ("side_modulus", C.c_double),
        ("has_rear_curve", C.c_int32),
        ("has_rear_modulus", C.c_int32),
        ("rear_curve", C.c_double * 2),
        ("rear_modulus", C.c_double),
        ("n_force_table", C.c_int32),
        ("n_compaction_map", C.c_int32),
        ("force_table", (C.c_double * 2) * MAX_TABLE),
        ("compaction_map", C.c_double * MAX_TABLE),
        ("has_crush_element_size", C.c_int32),
        ("has_bumper", C.c_int32),
        ("crush_element_size", C.c_double),
        ("bumper", C.c_double * 3),
        ("transverse_factor", C.c_double),
        ("has_crush_zone_mass", C.c_int32),
        ("has_rail_box", C.c_int32),
        ("crush_zone_mass", C.c_double),
        ("rail_box_width", C.c_double),
        ("rail_box_z", C.c_double * 2),
        ("rail_box_force_fraction", C.c_double),
    ]


class HeadOnOptions(C.Structure):
    _fields_ = [
        ("offset", C.c_double),
        ("clearance", C.c_double),
        ("element_size", C.c_double),
        ("end_time", C.c_double),
    ]


class BatchOptions(C.Structure):
    _fields_ = [
        ("threads_per_instance", C.c_int32),
        ("force_tile", C.c_int32),
        ("steps_per_graph", C.c_int32),
        ("host_threads", C.c_int32),
        ("contact_cell_scale", C.c_double),
        ("autotune", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("prepare_ms", C.c_double),
        ("upload_ms", C.c_double),
        ("device_ms", C.c_double),
        ("wall_ms", C.c_double),
        ("instances", C.c_int64),
        ("element_steps", C.c_int64),
        ("launches", C.c_int64),
        ("graph_replays", C.c_int64),
        ("pool_bytes", C.c_double),
        ("latency_p50_ms", C.c_double),
        ("latency_p99_ms", C.c_double),
        ("hc_element_steps", C.c_int64),
        ("k_end_ms", C.c_double),
        ("k_contact_ms", C.c_double),
        ("k_node_ms", C.c_double),
        ("k_force_hc_ms", C.c_double),
        ("k_force_el_ms", C.c_double),
        ("k_ms", C.c_double * 12),
    ]


class ServerOptions(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_double),
        ("max_resident", C.c_int32),
        ("steps_per_graph", C.c_int32),
        ("threads_per_instance", C.c_int32),
        ("force_tile", C.c_int32),
        ("contact_cell_scale", C.c_double),
        ("max_queued", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("admitted", C.c_int64),
        ("completed", C.c_int64),
        ("failed", C.c_int64),
        ("resident", C.c_int32),
        ("queued", C.c_int32),
        ("device_bytes_in_use", C.c_double),
        ("latency_p50_ms", C.c_double),
        ("latency_p99_ms", C.c_double),
        ("queue_p50_ms", C.c_double),
        ("busy_ms", C.c_double),
        ("element_steps", C.c_int64),
        ("replays", C.c_int64),
        ("captures", C.c_int64),
        ("admit_ms", C.c_double),
        ("lists_ms", C.c_double),
        ("replay_ms", C.c_double),
        ("retire_ms", C.c_double),
    ]


SUBMIT_FIELDS = 1


class PulseMetrics(C.Structure):
    _fields_ = [
        ("t_max_crush", C.c_double),
        ("max_crush", C.c_double),
        ("v_end", C.c_double),
        ("restitution", C.c_double),
        ("peak_accel", C.c_double),
    ]

LIB_PATH = Path(__file__).resolve().parent / "libcrushrs_amd.so"

# Every symbol the public headers declare, with (restype, argtypes).
SIGNATURES = {
    "crush_abi_version": (C.c_int32, []),
    "crush_last_error": (C.c_char_p, []),
    "crush_device_count": (C.c_int32, []),
    "crush_run": (C.c_int32, [C.POINTER(ModelView), C.POINTER(Results)]),
    "crush_run_noalloc": (C.c_int32, [C.POINTER(ModelView), C.POINTER(Results)]),
    "crush_debug_run_steps": (C.c_int32, [C.POINTER(ModelView), C.c_int32, f64p, f64p, f64p, i32p, u8p, f64p, i32p]),
    "crush_contact_node_count": (C.c_int32, [C.POINTER(ModelView), i32p]),
    "crush_output_capacity": (C.c_int32, [C.POINTER(ModelView), i32p, i32p]),
    "crush_batch_defaults": (None, [C.POINTER(BatchOptions)]),
    "crush_batch_create": (C.c_int32, [C.POINTER(C.POINTER(ModelView)), C.c_int32, C.POINTER(BatchOptions),
                                       C.POINTER(C.c_void_p)]),
    "crush_batch_run": (C.c_int32, [C.c_void_p, C.POINTER(BatchStats)]),
    "crush_batch_reset": (C.c_int32, [C.c_void_p]),
    "crush_batch_options_get": (C.c_int32, [C.c_void_p, C.POINTER(BatchOptions)]),
    "crush_batch_time_kernels": (C.c_int32, [C.c_void_p, C.c_int32]),
    "crush_batch_class_scales": (C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int32)]),
    "crush_batch_result": (C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(Results)]),
    "crush_batch_free": (None, [C.c_void_p]),
    "crush_server_defaults": (None, [C.POINTER(ServerOptions)]),
    "crush_server_create": (C.c_int32, [C.POINTER(ServerOptions), C.POINTER(C.c_void_p)]),
    "crush_server_submit": (C.c_int32, [C.c_void_p, C.POINTER(ModelView), C.c_int32, C.POINTER(C.c_int64)]),
    "crush_server_poll": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(Results), C.POINTER(C.c_int32)]),
    "crush_server_wait": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(Results), C.c_double]),
    "crush_server_stats_get": (C.c_int32, [C.c_void_p, C.POINTER(ServerStats)]),
    "crush_server_destroy": (None, [C.c_void_p]),
    "crush_headon_defaults": (None, [C.POINTER(HeadOnOptions)]),
    "crush_vehicle_load": (C.c_int32, [C.c_char_p, C.POINTER(Vehicle)]),
    "crush_vehicle_side_profile": (C.c_int32, [C.POINTER(Vehicle), C.POINTER(Vehicle)]),
    "crush_vehicle_rear_profile": (C.c_int32, [C.POINTER(Vehicle), C.POINTER(Vehicle)]),
    "crush_scenario_headon": (C.c_int32, [C.POINTER(Vehicle), C.POINTER(Vehicle), C.c_double, C.c_double,
                                          C.POINTER(HeadOnOptions), C.POINTER(C.c_void_p)]),
    "crush_scenario_crash": (C.c_int32, [C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    "crush_scenario_tbone": (C.c_int32, [C.POINTER(Vehicle), C.POINTER(Vehicle), C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    "crush_scenario_barrier": (C.c_int32, [C.POINTER(Vehicle), C.c_double, C.c_double, C.c_double,
                                           C.POINTER(C.c_void_p)]),
    "crush_model_view": (C.POINTER(ModelView), [C.c_void_p]),
    "crush_model_set_settings": (C.c_int32, [C.c_void_p, C.POINTER(Settings)]),
    "crush_model_node_set": (C.c_int32, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(i32p), i32p]),
    "crush_model_part_names": (C.c_int32, [C.c_void_p, C.c_char_p, C.c_int32]),
    "crush_model_free": (None, [C.c_void_p]),
    "crush_cfc_filter": (C.c_int32, [f64p, C.c_int32, C.c_double, C.c_double, f64p]),
    "crush_resample": (C.c_int32, [f64p, f64p, C.c_int32, C.c_double, C.c_double, C.c_double, f64p, C.c_int32,
                                   i32p]),
    "crush_integrate": (C.c_int32, [f64p, f64p, C.c_int32, C.c_double, f64p]),
    "crush_pulse_metrics_get": (C.c_int32, [f64p, f64p, C.c_int32, C.c_double, C.c_double,
                                            C.POINTER(PulseMetrics)]),
}

_lib = None


class CrushError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load libcrushrs_amd.so (built in-tree).  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built: run `python -m crushrs_amd.build`")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int) -> None:
    if code != CRUSH_OK:
        msg = load().crush_last_error()
        raise CrushError(code, msg.decode() if msg else "")
