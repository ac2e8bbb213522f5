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

from . import MPH, Model, renamed, vehicle_file

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
