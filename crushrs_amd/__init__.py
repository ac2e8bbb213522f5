"""MI355X crash-solver core: Python mirror of the crushrs solver API.

The reference's interface is `crushrs::run(&Model) -> Results`
(src/solver.rs:190) plus the scenario builders of src/vehicle.rs and
src/headon.rs.  Here the same calls go through the C ABI of
libcrushrs_amd.so (include/crushrs.h, include/crushrs_scenario.h); the
solve itself runs on the GPU.  Nothing in this package computes on the CPU
beyond model construction.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

from . import _abi
from ._abi import CrushError, check, load

MPH = 0.44704  # src/lib.rs:28
GROUND_CLEARANCE = 0.15  # src/vehicle.rs:499
TUNED_ELEMENT_SIZE = 0.3  # src/vehicle.rs:647
PULSE_END = 0.15  # src/vehicle.rs:1104

VEHICLE_DIR = Path(__file__).resolve().parent / "vehicles"


def vehicle(name: str) -> _abi.Vehicle:
    """A preset (neon, silverado, expedition, navigator, neon-pulse, ...)
    or a vehicle TOML path (src/bin/crushrs.rs:246-260)."""
    v = _abi.Vehicle()
    check(load().crush_vehicle_load(str(name).encode(), C.byref(v)))
    return v


def vehicle_file(stem: str) -> _abi.Vehicle:
    """One of the tuned vehicle TOMLs shipped with the package
    (neon_1996, silverado_2007, expedition_1999, navigator_1999)."""
    return vehicle(str(VEHICLE_DIR / f"{stem}.toml"))


def renamed(v: _abi.Vehicle, name: str) -> _abi.Vehicle:
    out = _abi.Vehicle()
    C.pointer(out)[0] = v
    out.name = name.encode()
    return out


def side_profile(v: _abi.Vehicle) -> _abi.Vehicle:
    out = _abi.Vehicle()
    check(load().crush_vehicle_side_profile(C.byref(v), C.byref(out)))
    return out


def rear_profile(v: _abi.Vehicle) -> _abi.Vehicle:
    out = _abi.Vehicle()
    check(load().crush_vehicle_rear_profile(C.byref(v), C.byref(out)))
    return out


def default_settings() -> _abi.Settings:
    """Settings::default (src/model.rs:134-149)."""
    s = _abi.Settings()
    s.end_time = 0.15
    s.dt_scale = 0.85
    s.time_step = 0.0
    s.hourglass = 0.1
    s.velocity_damping = 1.0
    s.mass_scaling = 0.25
    s.adaptive_check_steps = 5
    s.separation_steps = 0
    return s


class Model:
    """An owned, flattened model (the reference's `Model`)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self._lib = load()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.crush_model_free(h)
            self._h = None

    # ---- builders
    @classmethod
    def crash(cls, truck_mph: float = 30.0, car_mph: float = 30.0) -> "Model":
        """`crushrs crash truck_mph car_mph` (src/bin/crushrs.rs:403-432)."""
        h = C.c_void_p()
        check(load().crush_scenario_crash(truck_mph, car_mph, C.byref(h)))
        return cls(h)

    @classmethod
    def headon(cls, a: _abi.Vehicle, b: _abi.Vehicle, speed_a: float, speed_b: float, offset: float = 0.0,
               clearance: Optional[float] = GROUND_CLEARANCE, element_size: float = TUNED_ELEMENT_SIZE,
               end_time: float = PULSE_END) -> "Model":
        """HeadOn::build (src/headon.rs:40-66); speeds in m/s."""
        o = _abi.HeadOnOptions(offset, -1.0 if clearance is None else clearance, element_size, end_time)
        h = C.c_void_p()
        check(load().crush_scenario_headon(C.byref(a), C.byref(b), speed_a, speed_b, C.byref(o), C.byref(h)))
        return cls(h)

    @classmethod
    def tbone(cls, striking: _abi.Vehicle, struck: _abi.Vehicle, speed_a: float, speed_b: float,
              offset: float = 0.0, clearance: Optional[float] = GROUND_CLEARANCE,
              element_size: float = TUNED_ELEMENT_SIZE) -> "Model":
        """`crushrs tbone` generalised (src/bin/crushrs.rs:550-573)."""
        h = C.c_void_p()
        check(load().crush_scenario_tbone(C.byref(striking), C.byref(struck), speed_a, speed_b, offset,
                                          -1.0 if clearance is None else clearance, element_size, C.byref(h)))
        return cls(h)

    @classmethod
    def barrier(cls, v: _abi.Vehicle, element_size: float = TUNED_ELEMENT_SIZE, speed: float = 35 * MPH,
                end_time: float = PULSE_END) -> "Model":
        """barrier_model (src/vehicle.rs:951-973)."""
        h = C.c_void_p()
        check(load().crush_scenario_barrier(C.byref(v), element_size, speed, end_time, C.byref(h)))
        return cls(h)

    # ---- access
    @property
    def view(self) -> _abi.ModelView:
        return self._lib.crush_model_view(self._h).contents

    @property
    def ptr(self):
        return self._lib.crush_model_view(self._h)

    @property
    def settings(self) -> _abi.Settings:
        s = _abi.Settings()
        C.pointer(s)[0] = self.view.settings
        return s

    @settings.setter
    def settings(self, s: _abi.Settings) -> None:
        check(self._lib.crush_model_set_settings(self._h, C.byref(s)))

    def set(self, **kw) -> "Model":
        s = self.settings
        for k, v in kw.items():
            setattr(s, k, v)
        self.settings = s
        return self

    @property
    def n_nodes(self) -> int:
        return self.view.n_nodes

    @property
    def n_hexes(self) -> int:
        return self.view.n_hexes

    def nodes(self) -> np.ndarray:
        v = self.view
        return np.ctypeslib.as_array(v.nodes, shape=(v.n_nodes, 3)).copy()

    def hexes(self) -> np.ndarray:
        v = self.view
        return np.ctypeslib.as_array(v.hexes, shape=(v.n_hexes, 8)).copy()

    def initial_velocity(self) -> np.ndarray:
        v = self.view
        return np.ctypeslib.as_array(v.initial_velocity, shape=(v.n_nodes, 3)).copy()

    def node_set(self, name: str, face_set: bool = False) -> np.ndarray:
        p = _abi.i32p()
        n = C.c_int32()
        check(self._lib.crush_model_node_set(self._h, name.encode(), int(face_set), C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else np.zeros(0, np.int32)

    def part_names(self) -> list[str]:
        buf = C.create_string_buffer(1 << 16)
        check(self._lib.crush_model_part_names(self._h, buf, len(buf)))
        return buf.value.decode().split("\n")

    def contact_node_count(self) -> int:
        n = C.c_int32()
        check(self._lib.crush_contact_node_count(self.ptr, C.byref(n)))
        return n.value

    def output_capacity(self) -> tuple[int, int]:
        """(history samples, frames) a solve of this model can report."""
        if self.settings.history_steps <= 0 and self.settings.frame_steps <= 0:
            return 0, 0
        h, f = C.c_int32(), C.c_int32()
        check(self._lib.crush_output_capacity(self.ptr, C.byref(h), C.byref(f)))
        return h.value, f.value


@dataclass
class GroupResult:
    mass: float
    delta_v: np.ndarray
    crush_end: float
    crush_max: float
    pulse_delta_v: float = 0.0


@dataclass
class Results:
    """src/solver.rs:63-80, plus the per-vehicle outputs."""
    time: float
    steps: int
    eroded_elements: int
    wall_ms: float
    device_ms: float
    added_mass_max: float
    separated: bool
    displacement: np.ndarray
    velocity: np.ndarray
    masses: np.ndarray
    groups: list = field(default_factory=list)
    # accelerometer pulses [n_groups][samples], unfiltered local-x m/s^2 at
    # t = k * pulse_dt (None when the model records none)
    pulse: Optional[np.ndarray] = None
    pulse_dt: float = 0.0
    # histories (settings.history_steps > 0): sample times and steps [K],
    # part states [K][n_parts][PART_FIELDS] (mass, u xyz, v xyz, a xyz,
    # kinetic energy), accelerometer samples [K][min(n_groups, 8)][ACCEL_FIELDS]
    # (position, u, v, a xyz each, frame origin xyz, frame axes 3x3 rows)
    history_time: Optional[np.ndarray] = None
    history_step: Optional[np.ndarray] = None
    parts: Optional[np.ndarray] = None
    accel: Optional[np.ndarray] = None
    # frames (settings.frame_steps > 0, one-model path): times [F],
    # displacement [F][n_nodes][3], plastic strain [F][n_hexes], Cauchy
    # stress [F][n_hexes][6] (xx yy zz xy yz xz), eroded [F][n_hexes]
    frame_time: Optional[np.ndarray] = None
    frame_displacement: Optional[np.ndarray] = None
    frame_plastic_strain: Optional[np.ndarray] = None
    frame_stress: Optional[np.ndarray] = None
    frame_eroded: Optional[np.ndarray] = None

    def mean_velocity(self, nodes) -> tuple[float, np.ndarray]:
        """Results::mean_velocity / Model::mean_velocity (src/model.rs:281-295)."""
        nodes = np.asarray(nodes)
        m = self.masses[nodes]
        mt = m.sum()
        if mt <= 0:
            return 0.0, np.zeros(3)
        return float(mt), (m[:, None] * self.velocity[nodes]).sum(0) / mt


PULSE_CAPACITY = 8192  # samples per group kept in Results.pulse (0.8 s at 1e-4 s)


def alloc_results(n_nodes: int, n_groups: int, pulse_capacity: int = PULSE_CAPACITY, n_parts: int = 0,
                  history_capacity: int = 0, n_hexes: int = 0, frame_capacity: int = 0):
    u = np.zeros((n_nodes, 3))
    v = np.zeros((n_nodes, 3))
    m = np.zeros(n_nodes)
    g = (_abi.GroupResult * max(n_groups, 1))()
    r = _abi.Results()
    r.displacement = u.ctypes.data_as(_abi.f64p)
    r.velocity = v.ctypes.data_as(_abi.f64p)
    r.masses = m.ctypes.data_as(_abi.f64p)
    r.groups = C.cast(g, C.POINTER(_abi.GroupResult))
    p = np.zeros((max(min(n_groups, 8), 1), pulse_capacity))
    r.pulse = p.ctypes.data_as(_abi.f64p)
    r.pulse_capacity = pulse_capacity
    bufs = {"pulse": p}  # kept alive with the struct
    if history_capacity > 0:
        k = history_capacity
        bufs["history_time"] = np.zeros(k)
        bufs["history_step"] = np.zeros(k, np.int64)
        bufs["parts"] = np.zeros((k, max(n_parts, 1), _abi.PART_FIELDS))
        bufs["accel"] = np.zeros((k, max(min(n_groups, 8), 1), _abi.ACCEL_FIELDS))
        r.history_time = bufs["history_time"].ctypes.data_as(_abi.f64p)
        r.history_step = bufs["history_step"].ctypes.data_as(C.POINTER(C.c_int64))
        r.history = bufs["parts"].ctypes.data_as(_abi.f64p)
        r.accel_history = bufs["accel"].ctypes.data_as(_abi.f64p)
        r.history_capacity = k
    if frame_capacity > 0:
        f = frame_capacity
        bufs["frame_time"] = np.zeros(f)
        bufs["frame_displacement"] = np.zeros((f, n_nodes, 3), np.float32)
        bufs["frame_plastic_strain"] = np.zeros((f, n_hexes), np.float32)
        bufs["frame_stress"] = np.zeros((f, n_hexes, 6), np.float32)
        bufs["frame_eroded"] = np.zeros((f, n_hexes), np.uint8)
        r.frame_time = bufs["frame_time"].ctypes.data_as(_abi.f64p)
        r.frame_displacement = bufs["frame_displacement"].ctypes.data_as(C.POINTER(C.c_float))
        r.frame_plastic_strain = bufs["frame_plastic_strain"].ctypes.data_as(C.POINTER(C.c_float))
        r.frame_stress = bufs["frame_stress"].ctypes.data_as(C.POINTER(C.c_float))
        r.frame_eroded = bufs["frame_eroded"].ctypes.data_as(_abi.u8p)
        r.frame_capacity = f
    r.buffers = bufs
    return r, u, v, m, g


def alloc_model_results(model: "Model"):
    """alloc_results sized for everything a solve of `model` reports."""
    hc, fc = model.output_capacity()
    v = model.view
    return alloc_results(v.n_nodes, v.n_groups, PULSE_CAPACITY, v.n_parts, hc, v.n_hexes, fc)


def to_results(r, u, v, m, g, pulse_dt: float = 0.0) -> Results:
    groups = [GroupResult(g[i].mass, np.array(g[i].delta_v[:]), g[i].crush_end, g[i].crush_max, g[i].pulse_delta_v)
              for i in range(r.n_groups)]
    bufs = getattr(r, "buffers", {})
    pulse = None
    if "pulse" in bufs and r.pulse_samples > 0:
        ns = min(int(r.pulse_samples), r.pulse_capacity)
        pulse = bufs["pulse"][:min(r.n_groups, 8), :ns].copy()
    out = Results(r.time, int(r.steps), int(r.eroded_elements), r.wall_ms, r.device_ms, r.added_mass_max,
                  bool(r.separated), u, v, m, groups, pulse, pulse_dt)
    if "history_time" in bufs and r.history_samples > 0:
        k = min(int(r.history_samples), r.history_capacity)
        out.history_time = bufs["history_time"][:k].copy()
        out.history_step = bufs["history_step"][:k].copy()
        out.parts = bufs["parts"][:k].copy()
        out.accel = bufs["accel"][:k, :min(r.n_groups, 8)].copy()
    if "frame_time" in bufs and r.frame_samples > 0:
        f = min(int(r.frame_samples), r.frame_capacity)
        for name in ("frame_time", "frame_displacement", "frame_plastic_strain", "frame_stress", "frame_eroded"):
            setattr(out, name, bufs[name][:f].copy())
    return out


def run(model: Model) -> Results:
    """crushrs::run on the GPU (src/solver.rs:190-381)."""
    lib = load()
    r, u, v, m, g = alloc_model_results(model)
    check(lib.crush_run(model.ptr, C.byref(r)))
    return to_results(r, u, v, m, g, model.settings.pulse_dt)


def run_noalloc(model: Model) -> Results:
    """run() with no device memory allocated: the solve's device buffers are
    carved from the library's static device scratch (crush_run_noalloc)."""
    lib = load()
    r, u, v, m, g = alloc_model_results(model)
    check(lib.crush_run_noalloc(model.ptr, C.byref(r)))
    return to_results(r, u, v, m, g, model.settings.pulse_dt)


def debug_run_steps(model: Model, n_steps: int):
    """Per-step taps of the first n_steps on the GPU: f_int, f_ext, hit,
    eroded, dt (see crush_debug_run_steps)."""
    lib = load()
    n3 = 3 * model.n_nodes
    k = model.contact_node_count()
    fi = np.zeros((n_steps, n3))
    fx = np.zeros((n_steps, n3))
    hit = np.zeros((n_steps, max(k, 1)), np.int32)
    er = np.zeros((n_steps, model.n_hexes), np.uint8)
    dt = np.zeros(n_steps)
    done = C.c_int32()
    check(lib.crush_debug_run_steps(model.ptr, n_steps, fi.ctypes.data_as(_abi.f64p), fx.ctypes.data_as(_abi.f64p),
                                    None, hit.ctypes.data_as(_abi.i32p), er.ctypes.data_as(_abi.u8p),
                                    dt.ctypes.data_as(_abi.f64p), C.byref(done)))
    s = done.value
    return fi[:s], fx[:s], hit[:s, :k], er[:s], dt[:s]


def batch_options(**kw) -> _abi.BatchOptions:
    """crush_batch_defaults with fields overridden by keyword."""
    o = _abi.BatchOptions()
    load().crush_batch_defaults(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Batch:
    """Many independent models solved together on the GPU (the reference's
    par_iter over solver::run, src/vehicle.rs:1227-1233).  Results of each
    instance match crush_run on that model alone."""

    def __init__(self, models, options: Optional[_abi.BatchOptions] = None, **kw):
        self._lib = load()
        self.models = list(models)
        self._n_nodes = [m.n_nodes for m in self.models]
        self._n_groups = [m.view.n_groups for m in self.models]
        arr = (C.POINTER(_abi.ModelView) * len(self.models))(*[m.ptr for m in self.models])
        opt = options if options is not None else batch_options(**kw)
        self._h = C.c_void_p()
        check(self._lib.crush_batch_create(arr, len(self.models), C.byref(opt), C.byref(self._h)))
        self.stats = _abi.BatchStats()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.crush_batch_free(h)
            self._h = None

    def __len__(self) -> int:
        return len(self.models)

    def run(self) -> _abi.BatchStats:
        check(self._lib.crush_batch_run(self._h, C.byref(self.stats)))
        return self.stats

    def reset(self) -> None:
        check(self._lib.crush_batch_reset(self._h))

    def time_kernels(self, on: bool = True) -> None:
        check(self._lib.crush_batch_time_kernels(self._h, int(on)))

    @property
    def options(self) -> _abi.BatchOptions:
        o = _abi.BatchOptions()
        check(self._lib.crush_batch_options_get(self._h, C.byref(o)))
        return o

    def class_scales(self) -> list[tuple[int, float]]:
        """Autotuned contact grid cell scale per size class: (largest element
        count of the class, scale), classes ascending; empty when the batch
        was not autotuned."""
        scale, elems, k = (C.c_double * 16)(), (C.c_int32 * 16)(), C.c_int32()
        check(self._lib.crush_batch_class_scales(self._h, 16, scale, elems, C.byref(k)))
        return [(int(elems[i]), float(scale[i])) for i in range(min(k.value, 16))]

    def result(self, i: int, fields: bool = True) -> Results:
        mdl = self.models[i]
        hc = mdl.output_capacity()[0] if mdl.settings.history_steps > 0 else 0
        r, u, v, m, g = alloc_results(self._n_nodes[i], self._n_groups[i], PULSE_CAPACITY, mdl.view.n_parts, hc)
        if not fields:
            r.displacement = None
            r.velocity = None
        check(self._lib.crush_batch_result(self._h, i, C.byref(r)))
        return to_results(r, u, v, m, g, self.models[i].settings.pulse_dt)


def server_options(**kw) -> _abi.ServerOptions:
    """crush_server_defaults with fields overridden by keyword."""
    o = _abi.ServerOptions()
    load().crush_server_defaults(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Server:
    """Continuous-batching request loop on the GPU (crush_server_*): submit
    models from any thread, poll or wait for results."""

    def __init__(self, options: Optional[_abi.ServerOptions] = None, **kw):
        self._lib = load()
        opt = options if options is not None else server_options(**kw)
        self._h = C.c_void_p()
        check(self._lib.crush_server_create(C.byref(opt), C.byref(self._h)))
        self._shape = {}

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            self._lib.crush_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, model: Model, fields: bool = False) -> int:
        t = C.c_int64()
        check(self._lib.crush_server_submit(self._h, model.ptr, _abi.SUBMIT_FIELDS if fields else 0, C.byref(t)))
        hc = model.output_capacity()[0] if model.settings.history_steps > 0 else 0
        self._shape[t.value] = (model.n_nodes, model.view.n_groups, fields, model.settings.pulse_dt,
                                model.view.n_parts, hc)
        return t.value

    def _alloc(self, ticket):
        # an unknown ticket gets empty outputs; the library reports the error
        n_nodes, n_groups, fields, pulse_dt, n_parts, hc = self._shape.get(ticket, (0, 0, False, 0.0, 0, 0))
        r, u, v, m, g = alloc_results(n_nodes, n_groups, PULSE_CAPACITY, n_parts, hc)
        if not fields:
            r.displacement = None
            r.velocity = None
        r.pulse_dt_value = pulse_dt
        return r, u, v, m, g

    def poll(self, ticket: int) -> Optional[Results]:
        r, u, v, m, g = self._alloc(ticket)
        done = C.c_int32()
        check(self._lib.crush_server_poll(self._h, ticket, C.byref(r), C.byref(done)))
        if not done.value:
            return None
        self._shape.pop(ticket, None)
        return to_results(r, u, v, m, g, r.pulse_dt_value)

    def wait(self, ticket: int, timeout_ms: float = 0.0) -> Results:
        r, u, v, m, g = self._alloc(ticket)
        check(self._lib.crush_server_wait(self._h, ticket, C.byref(r), timeout_ms))
        self._shape.pop(ticket, None)
        return to_results(r, u, v, m, g, r.pulse_dt_value)

    def stats(self) -> _abi.ServerStats:
        st = _abi.ServerStats()
        check(self._lib.crush_server_stats_get(self._h, C.byref(st)))
        return st


def device_count() -> int:
    return int(load().crush_device_count())


__all__ = ["MPH", "Model", "Results", "run", "vehicle", "vehicle_file", "side_profile", "rear_profile",
           "default_settings", "debug_run_steps", "device_count", "CrushError", "renamed", "Batch", "batch_options", "Server",
           "server_options"]
