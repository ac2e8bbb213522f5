"""Seeded synthetic workloads built with the scenario builders.

The batched sweep draws head-on/offset pairs of the four tuned vehicle
TOMLs shipped with the reference (vehicles/*.toml), with speeds uniform in
10-40 mph per vehicle, lateral offset uniform in 0-0.6 m (HeadOn.offset,
src/headon.rs:23-24) and element size 0.3 or 0.2 m, each pair on the rigid
road 0.15 m below (src/headon.rs:46-49).
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import MPH, Model, rear_profile, renamed, side_profile, vehicle, vehicle_file

SWEEP_VEHICLES = ("neon_1996", "silverado_2007", "expedition_1999", "navigator_1999")


@dataclass(frozen=True)
class HeadOnCase:
    a: str
    b: str
    speed_a_mph: float
    speed_b_mph: float
    offset: float
    element_size: float
    end_time: float = 0.15

    def build(self, cache=None) -> Model:
        cache = {} if cache is None else cache
        va = cache.get(self.a) or vehicle_file(self.a)
        vb = cache.get(self.b) or vehicle_file(self.b)
        return Model.headon(renamed(va, f"{self.a}_a"), renamed(vb, f"{self.b}_b"), self.speed_a_mph * MPH,
                            self.speed_b_mph * MPH, offset=self.offset, element_size=self.element_size,
                            end_time=self.end_time)


def headon_sweep(n: int, seed: int = 0, sizes=(0.3, 0.2), end_time: float = 0.15) -> list[HeadOnCase]:
    """n seeded head-on/offset cases."""
    rng = np.random.default_rng(seed)
    pairs = rng.integers(0, len(SWEEP_VEHICLES), size=(n, 2))
    speeds = rng.uniform(10.0, 40.0, size=(n, 2))
    offsets = rng.uniform(0.0, 0.6, size=n)
    h = rng.choice(np.asarray(sizes, dtype=float), size=n)
    return [HeadOnCase(SWEEP_VEHICLES[pairs[i, 0]], SWEEP_VEHICLES[pairs[i, 1]], float(speeds[i, 0]),
                       float(speeds[i, 1]), float(offsets[i]), float(h[i]), end_time) for i in range(n)]


def build_models(cases, threads: int = 16) -> list[Model]:
    """Build the models of `cases` on host threads (the builders run in
    native code and release the interpreter lock)."""
    cache = {name: vehicle_file(name) for name in {c.a for c in cases} | {c.b for c in cases}}
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda c: c.build(cache), cases))


def honeycomb_count(model: Model) -> int:
    """Number of honeycomb elements of a model."""
    v = model.view
    parts = np.ctypeslib.as_array(v.hex_part, shape=(v.n_hexes,))
    kinds = np.array([v.materials[p].model for p in range(v.n_parts)], dtype=np.int32)
    return int(np.count_nonzero(kinds[parts] == 1))


# Rear crush coefficients for vehicles whose TOML has no [rear] section:
# the front curve scaled by this factor (the tuned TOMLs carry none, and
# rear_profile needs one, src/vehicle.rs:271-316).
SYNTHETIC_REAR_FACTOR = 0.7


def with_rear(v):
    """A copy of vehicle v with synthetic rear coefficients if it has none."""
    out = renamed(v, v.name.decode())
    if not out.has_rear_curve:
        out.has_rear_curve = 1
        out.rear_curve[0] = SYNTHETIC_REAR_FACTOR * v.curve[0]
        out.rear_curve[1] = SYNTHETIC_REAR_FACTOR * v.curve[1]
    return out


@dataclass(frozen=True)
class ServeCase:
    """One delta-v request: kind 'headon' (with lateral offset), 'tbone'
    (front of a into the side profile of b) or 'rear' (front of a into the
    rear profile of b, b moving the same way more slowly)."""
    kind: str
    a: str
    b: str
    speed_a_mph: float
    speed_b_mph: float
    offset: float
    element_size: float
    separation_steps: int = 20

    def build(self, cache=None) -> Model:
        cache = {} if cache is None else cache
        va = renamed(cache.get(self.a) or vehicle_file(self.a), f"{self.a}_a")
        vb = renamed(cache.get(self.b) or vehicle_file(self.b), f"{self.b}_b")
        h = self.element_size
        if self.kind == "headon":
            m = Model.headon(va, vb, self.speed_a_mph * MPH, self.speed_b_mph * MPH, offset=self.offset, element_size=h)
        elif self.kind == "tbone":
            m = Model.tbone(va, side_profile(vb), self.speed_a_mph * MPH, 0.0, offset=self.offset, element_size=h)
        else:
            m = Model.headon(va, rear_profile(with_rear(vb)), self.speed_a_mph * MPH, -self.speed_b_mph * MPH,
                             element_size=h)
        return m.set(separation_steps=self.separation_steps)


def serve_mix(n: int, seed: int = 0, sizes=(0.3, 0.15)) -> list[ServeCase]:
    """n seeded requests: 60 % head-on/offset, 25 % T-bone, 15 % rear-end;
    speeds uniform in 5-45 mph (rear-end: the struck vehicle at a random
    fraction of the striking speed), element size from `sizes`."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        u = rng.random()
        kind = "headon" if u < 0.60 else ("tbone" if u < 0.85 else "rear")
        a, b = (SWEEP_VEHICLES[i] for i in rng.integers(0, len(SWEEP_VEHICLES), size=2))
        sa = float(rng.uniform(5.0, 45.0))
        sb = float(rng.uniform(5.0, 45.0)) if kind == "headon" else (float(rng.uniform(0.0, 0.8)) * sa
                                                                     if kind == "rear" else 0.0)
        off = float(rng.uniform(0.0, 0.6)) if kind == "headon" else (float(rng.uniform(-0.5, 0.5))
                                                                      if kind == "tbone" else 0.0)
        out.append(ServeCase(kind, a, b, sa, sb, off, float(rng.choice(np.asarray(sizes, float)))))
    return out


# ---------------------------------------------------------------- large pair

# Crush distances (m) of the 7-knot force tables given to the two presets of
# the large single instance.  The forces are the presets' own bilinear
# curves, F = yield_force + stiffness * x (src/vehicle.rs:621-633), sampled
# at these distances; Vehicle::crush_material_axial then turns each table
# into a yield-vs-compaction curve and adds the lock-up knot
# [1.4, sigma(1.4)] with slope + E (src/vehicle.rs:478-489, 502-511), for 8
# knots in all.
LARGE_KNOTS_M = (0.0, 0.1, 0.2, 0.35, 0.5, 0.7, 0.9)


def tabulated(v, knots_m=LARGE_KNOTS_M):
    """A copy of vehicle v whose bilinear crush curve is given as a force
    table sampled at knots_m."""
    out = renamed(v, v.name.decode())
    out.n_force_table = len(knots_m)
    out.n_compaction_map = 0
    for i, x in enumerate(knots_m):
        out.force_table[i][0] = x
        out.force_table[i][1] = v.curve[0] + v.curve[1] * x
    return out


def large_pair(element_size: float = 0.075, speed_mph: float = 30.0, end_time: float = 0.15) -> Model:
    """The large single instance: the `crash` presets (Silverado 2622 kg vs
    Neon 1354 kg) remeshed uniformly at element_size, head-on at +-speed,
    no road, mass_scaling 0.25, each crush curve an 8-knot table with
    lock-up.  At 0.075 m: 76,662 hexes (78x27x25 + 58x23x18) and 84,416
    nodes."""
    a = tabulated(vehicle("silverado"))
    b = renamed(tabulated(vehicle("neon")), "neon")
    m = Model.headon(a, b, speed_mph * MPH, speed_mph * MPH, clearance=None, element_size=element_size,
                     end_time=end_time)
    return m.set(mass_scaling=0.25)


# ---------------------------------------------------------- calibration set

CALIB_VEHICLES = SWEEP_VEHICLES


@dataclass(frozen=True)
class BarrierCase:
    """One rigid-barrier test (barrier_model, src/vehicle.rs:951-973) of a
    vehicle whose crush curve is perturbed the way a calibration driver
    probes it: the whole curve scaled by yield_scale, each force-table knot
    by its own factor, and the rail box (vehicles with a crush zone and a
    rail box) given force_fraction and width."""
    vehicle: str
    yield_scale: float
    knot_scales: tuple
    rail_fraction: float
    rail_width: float
    element_size: float = 0.3
    speed_mph: float = 35.0
    end_time: float = 0.15

    def build(self, cache=None) -> Model:
        cache = {} if cache is None else cache
        v = renamed(cache.get(self.vehicle) or vehicle_file(self.vehicle), self.vehicle)
        if v.n_force_table > 0:
            for i in range(v.n_force_table):
                v.force_table[i][1] *= self.yield_scale * self.knot_scales[i]
        else:
            v.curve[0] *= self.yield_scale
        if v.has_rail_box:
            v.rail_box_force_fraction = self.rail_fraction
            v.rail_box_width = self.rail_width
        return Model.barrier(v, element_size=self.element_size, speed=self.speed_mph * MPH, end_time=self.end_time)


def calibration_set(n: int, seed: int = 0) -> list[BarrierCase]:
    """n seeded barrier tests at 35 mph, 0.3 m elements: vehicle drawn from
    the four tuned TOMLs, yield scale log-uniform in [0.8, 1.25], knots
    uniform in [0.9, 1.1], rail-box force_fraction U(0.5, 0.8) and width
    U(0.4, 0.8) m."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        name = CALIB_VEHICLES[int(rng.integers(0, len(CALIB_VEHICLES)))]
        ys = float(np.exp(rng.uniform(np.log(0.8), np.log(1.25))))
        ks = tuple(float(x) for x in rng.uniform(0.9, 1.1, size=8))
        out.append(BarrierCase(name, ys, ks, float(rng.uniform(0.5, 0.8)), float(rng.uniform(0.4, 0.8))))
    return out


def build_cases(cases, threads: int = 16) -> list[Model]:
    """Build any list of cases with a .build(cache) method on host threads."""
    names = set()
    for c in cases:
        for attr in ("a", "b", "vehicle"):
            if hasattr(c, attr):
                names.add(getattr(c, attr))
    cache = {name: vehicle_file(name) for name in names}
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda c: c.build(cache), cases))
