"""Calibration drivers as clients of the batched solver.

The reference calibrates vehicles by running many independent solves: the
pulse pattern search of `calibrate_pulse` (src/vehicle.rs:1127-1258) runs
the 18 probes of a round side by side on host threads, and the rail-box
differential evolution of `calibrate_rail_boxes` (src/headon.rs:124-250)
does the same with each generation's trial population.  Here each round (or
generation) is one `Batch` on the GPU: the probe vehicles become barrier (or
head-on) models, one instance each, solved together; the pulses come back
from the device and the search logic below runs on the host, as the
reference's does.

The search needs nothing but a way to score a list of candidates, so both
drivers take an optional `solve` (models -> results) to run the same search
through another path, e.g. one `run` per model.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import GROUND_CLEARANCE, PULSE_END, TUNED_ELEMENT_SIZE, Batch, Model, _abi
from .signal import PULSE_DT, Pulse, PulseComparison

# One knot of the reference's table is the densification (src/vehicle.rs:1148).
N_KNOTS = _abi.MAX_TABLE - 1
MODULUS_RANGE = (2e6, 500e6)  # crush-zone modulus clamp (src/vehicle.rs:1209)
BODY_MODULUS_RANGE = (2e7, 5e9)  # body modulus clamp (src/vehicle.rs:1211)


def copy_vehicle(v: _abi.Vehicle) -> _abi.Vehicle:
    out = _abi.Vehicle()
    C.pointer(out)[0] = v
    return out


def force_table(v: _abi.Vehicle) -> np.ndarray:
    return np.array([[v.force_table[i][0], v.force_table[i][1]] for i in range(v.n_force_table)]).reshape(-1, 2)


def set_force_table(v: _abi.Vehicle, table) -> None:
    table = np.asarray(table, dtype=float).reshape(-1, 2)
    if len(table) > _abi.MAX_TABLE:
        raise ValueError(f"force table has {len(table)} knots, at most {_abi.MAX_TABLE}")
    v.n_force_table = len(table)
    for i, (x, f) in enumerate(table):
        v.force_table[i][0] = x
        v.force_table[i][1] = f


# ---------------------------------------------------------------- pulse metrics

def force_crush(p: Pulse, mass: float, v0: float) -> np.ndarray:
    """The (crush, force) samples of the loading phase, F = -m*a against the
    integrated displacement up to maximum crush (PulseMetrics::new,
    src/signal.rs:232-240): the first sample of the largest positive
    displacement ends the phase, sample 0 when none is positive."""
    x = p.displacement(v0)
    i_max, best = 0, 0.0
    for i, xi in enumerate(x):
        if xi > best:
            i_max, best = i, xi
    return np.stack([x[:i_max + 1], -mass * p.accel[:i_max + 1]], axis=1)


def mean_force(fc: np.ndarray, x0: float, x1: float) -> Optional[float]:
    """Mean force over the loading-phase samples with x0 <= crush <= x1:
    trapezoid energy over the span those samples cover; None for fewer than
    two samples or an empty span (PulseMetrics::mean_force,
    src/signal.rs:274-291)."""
    pts = fc[(fc[:, 0] >= x0) & (fc[:, 0] <= x1)]
    if len(pts) < 2:
        return None
    e = float(np.sum(0.5 * (pts[:-1, 1] + pts[1:, 1]) * (pts[1:, 0] - pts[:-1, 0])))
    span = pts[-1, 0] - pts[0, 0]
    return e / span if span > 0.0 else None


def kw(fc: np.ndarray, x0: float, x1: float) -> Optional[float]:
    """Energy-equivalent stiffness 2E/(x1^2 - x0^2) over a crush window;
    None when the samples stop short of x1 - 5 cm (PulseMetrics::kw,
    src/signal.rs:293-303)."""
    pts = fc[(fc[:, 0] >= x0) & (fc[:, 0] <= x1)]
    if len(pts) < 2 or pts[-1, 0] < x1 - 0.05:
        return None
    e = float(np.sum(0.5 * (pts[:-1, 1] + pts[1:, 1]) * (pts[1:, 0] - pts[:-1, 0])))
    return 2.0 * e / (x1 * x1 - x0 * x0)


def pulse_objective(cmp: PulseComparison) -> float:
    """Velocity-history RMS error (m/s) plus half the end-velocity error plus
    a tenth of the CFC 60 acceleration RMS error in g (pulse_objective,
    src/vehicle.rs:1117-1123)."""
    v_rms = math.sqrt(float(np.mean((cmp.v_meas - cmp.v_sim) ** 2)))
    v_end = abs(float(cmp.v_sim[-1] - cmp.v_meas[-1]))
    return v_rms + 0.5 * v_end + 0.1 * cmp.rms_accel_error() / 9.81


# ---------------------------------------------------------------- solving

def solve_batched(models: Sequence[Model], **options) -> list:
    """All models as one batch on the GPU; results without nodal fields."""
    b = Batch(models, **options)
    b.run()
    return [b.result(i, fields=False) for i in range(len(models))]


def _finite(j: float) -> float:
    return j if math.isfinite(j) else math.inf


# ---------------------------------------------------------------- calibrate_pulse

@dataclass
class PulseFit:
    """What calibrate_pulse returns: the tuned vehicle, its comparison with
    the measured pulse and objective, and the search's record."""
    vehicle: _abi.Vehicle
    comparison: PulseComparison
    objective: float
    evaluations: int
    rounds: int
    batches: int
    start_objective: float = math.inf
    # one entry per round: (round, objective after it, step after it,
    # accepted probe as (parameter, direction) or None)
    history: list = field(default_factory=list)


def initial_force_table(meas: Pulse, mass: float, speed: float) -> np.ndarray:
    """N_KNOTS knots spread over 0.95 of the measured dynamic crush, each at
    the mean measured force over +-half a spacing around it (at least 1 kN),
    non-decreasing from the second knot on (src/vehicle.rs:1145-1163)."""
    fc = force_crush(meas, mass, speed)
    x_max = float(fc[-1, 0]) if len(fc) else 0.0
    spacing = x_max * 0.95 / (N_KNOTS - 1)
    knots = np.arange(N_KNOTS) * spacing
    w = 0.5 * spacing
    targets = []
    for x in knots:
        f = mean_force(fc, max(x - w, 0.0), min(x + w, x_max))
        targets.append(max(f if f is not None else 0.0, 1e3))
    for i in range(2, len(targets)):
        targets[i] = max(targets[i], targets[i - 1])
    return np.stack([knots, np.array(targets)], axis=1)


def pulse_candidate(base: _abi.Vehicle, param: int, factor: float) -> Optional[_abi.Vehicle]:
    """`base` with one parameter scaled by `factor`: a knot force (the table
    then raised to non-decreasing from the second knot), the crush-zone
    modulus or the body modulus (each clamped); None when a knot move
    leaves the table as it was (src/vehicle.rs:1193-1214)."""
    cand = copy_vehicle(base)
    if param < N_KNOTS:
        t = force_table(base)
        t0 = t.copy()
        t[param, 1] *= factor
        for i in range(2, len(t)):
            t[i, 1] = max(t[i, 1], t[i - 1, 1])
        if np.array_equal(t, t0):
            return None
        set_force_table(cand, t)
    elif param == N_KNOTS:
        cand.modulus = min(max(cand.modulus * factor, MODULUS_RANGE[0]), MODULUS_RANGE[1])
    else:
        cand.body_modulus = min(max(cand.body_modulus * factor, BODY_MODULUS_RANGE[0]), BODY_MODULUS_RANGE[1])
    return cand


def barrier_pulse_scorer(meas: Pulse, mass: float, speed: float, element_size: float,
                         solve: Callable[[Sequence[Model]], list]) -> Callable:
    """Scores candidate vehicles by run_barrier_pulse (src/vehicle.rs:1106-1115)
    against `meas`: each becomes a rigid-barrier model with its rear-seat
    accelerometer, all are solved by one `solve` call, and each CFC 60 pulse
    is compared with the measured one."""
    def score(vehicles):
        models = [Model.barrier(v, element_size=element_size, speed=speed, end_time=PULSE_END) for v in vehicles]
        out = []
        for r in solve(models):
            sim = Pulse.from_results(r, 0).filtered(60.0)
            cmp = PulseComparison.new(meas, sim, mass, speed)
            out.append((_finite(pulse_objective(cmp)), cmp))
        return out
    return score


def calibrate_pulse(v: _abi.Vehicle, measured: Pulse, speed: float, element_size: float = TUNED_ELEMENT_SIZE,
                    rounds: int = 40, solve: Optional[Callable] = None, score: Optional[Callable] = None,
                    log: Optional[Callable[[str], None]] = None) -> PulseFit:
    """Fit the tabulated force-crush curve, crush-zone modulus and body
    modulus of `v` so its simulated rigid-barrier pulse reproduces
    `measured` (calibrate_pulse, src/vehicle.rs:1144-1258).

    A vehicle whose table does not have N_KNOTS knots starts from the
    measured force-crush curve (initial_force_table, compaction map
    cleared).  Each round probes every parameter up and down by exp(+-step)
    in log space (step 0.3 at first), keeps the best probe when it lowers
    pulse_objective, else halves the step; the search ends after `rounds`
    rounds or when the step falls below 0.04.  Each round's probes are one
    batch; the starting vehicle's solve rides in the first round's batch.

    `score` (list of vehicles -> list of (objective, comparison)) replaces
    the barrier solves altogether; `solve` (list of models -> results)
    replaces only the batched GPU solve."""
    meas = measured.filtered(60.0).resampled(PULSE_DT).truncated(PULSE_END)
    if score is None:
        score = barrier_pulse_scorer(meas, v.mass, speed, element_size, solve or solve_batched)
    tuned = copy_vehicle(v)
    if tuned.n_force_table != N_KNOTS:
        set_force_table(tuned, initial_force_table(meas, v.mass, speed))
        tuned.n_compaction_map = 0

    def probes_of(base, step):
        out = []
        for param in range(N_KNOTS + 2):
            for d in (1.0, -1.0):
                c = pulse_candidate(base, param, math.exp(d * step))
                if c is not None:
                    out.append((param, d, c))
        return out

    step = 0.3
    evaluations = batches = 0
    best_j, best_cmp = math.inf, None
    started = False
    history = []
    rounds_run = 0
    for rnd in range(rounds):
        probes = probes_of(tuned, step)
        first, started = not started, True
        cands = ([tuned] if first else []) + [c for _, _, c in probes]
        scored = score(cands)
        evaluations += len(cands)
        batches += 1
        if first:
            best_j, best_cmp = scored[0]
            start_j = best_j
            scored = scored[1:]
            if log:
                log(f"pulse calibrate {tuned.name.decode()} start: J {best_j:.4f}")
        rounds_run = rnd + 1
        accepted = None
        if scored:
            i = int(np.argmin([j for j, _ in scored]))
            if scored[i][0] < best_j:
                param, d, cand = probes[i]
                best_j, best_cmp = scored[i]
                tuned = cand
                accepted = (param, d)
        if accepted is None:
            step *= 0.5
        history.append((rnd, best_j, step, accepted))
        if log:
            log(f"pulse calibrate round {rnd}: J {best_j:.4f} step {step:.3f} "
                + (f"p{accepted[0]}{'+' if accepted[1] > 0 else '-'}" if accepted else "no improvement"))
        if accepted is None and step < 0.04:
            break
    if not started:  # rounds == 0: only the starting vehicle
        (best_j, best_cmp), = score([tuned])
        start_j = best_j
        evaluations += 1
        batches += 1
    return PulseFit(tuned, best_cmp, best_j, evaluations, rounds_run, batches, start_j, history)


# ---------------------------------------------------------------- calibrate_rail_boxes

class XorShift64Star:
    """xorshift64* (shifts 12, 25, 27; multiplier 0x2545F4914F6CDD1D), the
    reference's deterministic RNG for calibrations (src/headon.rs:104-122)."""
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        x = self.state
        x ^= x >> 12
        x ^= (x << 25) & self.MASK
        x ^= x >> 27
        self.state = x
        return (x * 0x2545F4914F6CDD1D) & self.MASK

    def uniform(self) -> float:
        return (self.next_u64() >> 11) / float(1 << 53)

    def below(self, n: int) -> int:
        return int(self.uniform() * n) % n


def rail_box_from_unit(v: _abi.Vehicle, u) -> tuple[float, tuple[float, float], float]:
    """A point of the unit cube as a rail box (width, (z0, z1), force
    fraction): 25-95 % of the width, centred at 15-50 % of the height,
    15-60 % of the height tall, 50-95 % of the force (src/headon.rs:80-94)."""
    w, hgt = v.size[1], v.size[2]
    width = (0.25 + 0.7 * u[0]) * w
    zc = (0.15 + 0.35 * u[1]) * hgt
    dz = (0.15 + 0.45 * u[2]) * hgt
    return width, (max(zc - 0.5 * dz, 0.0), min(zc + 0.5 * dz, hgt)), 0.5 + 0.45 * u[3]


def unit_from_rail_box(v: _abi.Vehicle) -> list:
    """The vehicle's rail box as a point of the unit cube, clamped
    (src/headon.rs:96-101)."""
    w, hgt = v.size[1], v.size[2]
    zc = 0.5 * (v.rail_box_z[0] + v.rail_box_z[1]) / hgt
    dz = (v.rail_box_z[1] - v.rail_box_z[0]) / hgt
    c = lambda x: min(max(x, 0.0), 1.0)  # noqa: E731
    return [c((v.rail_box_width / w - 0.25) / 0.7), c((zc - 0.15) / 0.35), c((dz - 0.15) / 0.45),
            c((v.rail_box_force_fraction - 0.5) / 0.45)]


def with_rail_box(v: _abi.Vehicle, u) -> _abi.Vehicle:
    out = copy_vehicle(v)
    width, (z0, z1), frac = rail_box_from_unit(v, u)
    out.has_rail_box = 1
    out.rail_box_width = width
    out.rail_box_z[0] = z0
    out.rail_box_z[1] = z1
    out.rail_box_force_fraction = frac
    return out


def with_default_rail_box(v: _abi.Vehicle) -> _abi.Vehicle:
    """A vehicle without a rail box gets one half its width wide, 20-60 %
    of its height, carrying 70 % of the force (src/headon.rs:152-158)."""
    out = copy_vehicle(v)
    if not out.has_rail_box:
        out.has_rail_box = 1
        out.rail_box_width = 0.5 * v.size[1]
        out.rail_box_z[0] = 0.2 * v.size[2]
        out.rail_box_z[1] = 0.6 * v.size[2]
        out.rail_box_force_fraction = 0.7
    return out


@dataclass
class HeadOnSetup:
    """A frontal vehicle-to-vehicle impact (HeadOn, src/headon.rs:16-38):
    A heads +x at speed_a, B heads -x at speed_b (m/s)."""
    speed_a: float
    speed_b: float
    offset: float = 0.0
    clearance: Optional[float] = GROUND_CLEARANCE
    element_size: float = TUNED_ELEMENT_SIZE
    end_time: float = PULSE_END

    def model(self, a: _abi.Vehicle, b: _abi.Vehicle) -> Model:
        return Model.headon(a, b, self.speed_a, self.speed_b, offset=self.offset, clearance=self.clearance,
                            element_size=self.element_size, end_time=self.end_time)

    @staticmethod
    def pulses(results) -> tuple[Pulse, Pulse]:
        """The two rear-seat pulses, CFC 60, each vehicle's forward
        direction positive (HeadOn::run, src/headon.rs:68-78)."""
        pa = Pulse.from_results(results, 0).filtered(60.0)
        pb = Pulse.from_results(results, 1).negated().filtered(60.0)
        return pa, pb


@dataclass
class RailBoxFit:
    """What calibrate_rail_boxes returns (RailBoxFit, src/headon.rs:124-133)."""
    a: _abi.Vehicle
    b: _abi.Vehicle
    objective: float
    cmp_a: PulseComparison
    cmp_b: PulseComparison
    evaluations: int
    batches: int
    # best and mean objective after the initial population and each generation
    history: list = field(default_factory=list)


def calibrate_rail_boxes(a: _abi.Vehicle, b: _abi.Vehicle, setup: HeadOnSetup, measured_a: Pulse, measured_b: Pulse,
                         population: int, generations: int, solve: Optional[Callable] = None,
                         score: Optional[Callable] = None, log: Optional[Callable[[str], None]] = None) -> RailBoxFit:
    """Fit both vehicles' rail boxes to a measured car-to-car test
    (calibrate_rail_boxes, src/headon.rs:135-250): the eight unit-cube
    parameters are searched by synchronous differential evolution
    (rand/1/bin, F 0.7, CR 0.9, population max(population, 4), seeded with
    the vehicles' current boxes, the rest drawn from xorshift64* seeded
    0x9E3779B97F4A7C15); the score of a member is the sum of the two
    vehicles' pulse_objective.  The initial population and each
    generation's trials are one batch of head-on solves.

    `score` (list of unit-cube points -> list of (objective, cmp_a, cmp_b))
    replaces the head-on solves altogether; `solve` (list of models ->
    results) replaces only the batched GPU solve."""
    a0, b0 = with_default_rail_box(a), with_default_rail_box(b)
    dims = 8

    def with_u(u):
        return with_rail_box(a0, u[:4]), with_rail_box(b0, u[4:])

    if score is None:
        end = setup.end_time
        meas_a = measured_a.filtered(60.0).resampled(PULSE_DT).truncated(end)
        meas_b = measured_b.filtered(60.0).resampled(PULSE_DT).truncated(end)
        run = solve or solve_batched

        def score(us):
            pairs = [with_u(u) for u in us]
            models = [setup.model(va, vb) for va, vb in pairs]
            out = []
            for r in run(models):
                pa, pb = setup.pulses(r)
                ca_ = PulseComparison.new(meas_a, pa, a.mass, setup.speed_a)
                cb_ = PulseComparison.new(meas_b, pb, b.mass, setup.speed_b)
                out.append((_finite(pulse_objective(ca_) + pulse_objective(cb_)), ca_, cb_))
            return out

    rng = XorShift64Star(0x9E3779B97F4A7C15)
    np_ = max(population, 4)
    pop = [unit_from_rail_box(a0) + unit_from_rail_box(b0)]
    while len(pop) < np_:
        pop.append([rng.uniform() for _ in range(dims)])
    scored = score(pop)
    evaluations, batches = len(pop), 1
    history = []

    def best_index():
        return int(np.argmin([s[0] for s in scored]))

    best = best_index()
    history.append((scored[best][0], float(np.mean([s[0] for s in scored]))))
    if log:
        log(f"rail-box calibrate: initial population of {np_}: best J {scored[best][0]:.4f}")
    f_weight, cr = 0.7, 0.9
    for gen in range(generations):
        trials = []
        for i in range(np_):
            while True:
                r = [rng.below(np_), rng.below(np_), rng.below(np_)]
                if i not in r and r[0] != r[1] and r[0] != r[2] and r[1] != r[2]:
                    break
            r1, r2, r3 = r
            j_rand = rng.below(dims)
            t = []
            for j in range(dims):
                if j == j_rand or rng.uniform() < cr:
                    x = pop[r1][j] + f_weight * (pop[r2][j] - pop[r3][j])
                    x = -x if x < 0.0 else (2.0 - x if x > 1.0 else x)  # reflect into the cube
                    t.append(min(max(x, 0.0), 1.0))
                else:
                    t.append(pop[i][j])
            trials.append(t)
        results = score(trials)
        evaluations += len(trials)
        batches += 1
        for i, (t, s) in enumerate(zip(trials, results)):
            if s[0] <= scored[i][0]:
                pop[i] = t
                scored[i] = s
        best = best_index()
        history.append((scored[best][0], float(np.mean([s[0] for s in scored]))))
        if log:
            log(f"rail-box calibrate: generation {gen} ({evaluations} runs): best J {scored[best][0]:.4f}, "
                f"mean {history[-1][1]:.4f}")
    va, vb = with_u(pop[best])
    j, cmp_a, cmp_b = scored[best]
    return RailBoxFit(va, vb, j, cmp_a, cmp_b, evaluations, batches, history)


# ---------------------------------------------------------------- barrier metrics and calibrate

BARRIER_END = 0.14  # end time of calibrate's barrier runs (src/vehicle.rs:1017)


@dataclass
class BarrierMetrics:
    """BarrierMetrics (src/vehicle.rs:866-881): energy-equivalent stiffness
    over the crush window (KW400 for 25-400 mm), maximum crush, peak and mean
    force of the loading phase, restitution, delta-v, and the (crush, force)
    samples of the loading phase."""
    kw400: float
    max_crush: float
    peak_force: float
    mean_force: float
    restitution: float
    delta_v: float
    curve: np.ndarray


def vehicle_parts(model: Model, name: str) -> list[int]:
    """The parts a vehicle was built as: <name>, <name>_crush, <name>_rails,
    <name>_bumper, <name>_bumper_rails, those that exist
    (src/vehicle.rs:890-891)."""
    names = model.part_names()
    out = [names.index(name + s) for s in ("", "_crush", "_rails", "_bumper", "_bumper_rails") if name + s in names]
    if not out:
        raise ValueError(f"no part '{name}'")
    return out


def _energy_at(curve: np.ndarray, d: float) -> float:
    e = 0.0
    for (c0, f0), (c1, f1) in zip(curve[:-1], curve[1:]):
        if c1 <= c0:
            continue
        if c1 <= d:
            e += 0.5 * (f0 + f1) * (c1 - c0)
        elif c0 < d:
            f_d = f0 + (f1 - f0) * (d - c0) / (c1 - c0)
            e += 0.5 * (f0 + f_d) * (d - c0)
            break
        else:
            break
    return e


def barrier_metrics(model: Model, results, name: str, axis: int = 0, window=(0.025, 0.4)) -> BarrierMetrics:
    """barrier_metrics_window (src/vehicle.rs:884-947) from the part history
    of a barrier solve (settings.history_steps > 0): the vehicle's
    mass-weighted mean velocity and displacement along `axis`, force
    -m dv/dt by central differences smoothed (1, 2, 1) / 4, contact where
    the force first exceeds 2 % of its peak, crush from there, the loading
    phase up to the last maximum of crush."""
    if results.parts is None:
        raise ValueError("the solve recorded no history (settings.history_steps = 0)")
    parts = vehicle_parts(model, name)
    P = results.parts[:, parts, :]
    m = P[:, :, 0].sum(axis=1)
    v = (P[:, :, 0] * P[:, :, 4 + axis]).sum(axis=1) / m
    u = (P[:, :, 0] * P[:, :, 1 + axis]).sum(axis=1) / m
    t = results.history_time
    n = len(t)
    if n <= 3:
        raise ValueError(f"no history for '{name}'")
    v0 = v[0]
    sign = math.copysign(1.0, v0)
    force = np.zeros(n)
    for i in range(1, n - 1):
        force[i] = -sign * m[i - 1] * (v[i + 1] - v[i - 1]) / (t[i + 1] - t[i - 1])
    sm = force.copy()
    for i in range(1, n - 1):
        sm[i] = (force[i - 1] + 2.0 * force[i] + force[i + 1]) / 4.0
    force = sm
    peak = max(0.0, float(force.max()))
    above = np.nonzero(force > 0.02 * peak)[0]
    contact = int(above[0]) if len(above) else 0
    crush = sign * (u - u[contact])
    i_max = n - 1 - int(np.argmax(crush[::-1]))
    max_crush = float(crush[i_max])
    curve = np.stack([crush[contact:i_max + 1], force[contact:i_max + 1]], axis=1)
    w0, w1 = window
    kw = 2.0 * (_energy_at(curve, w1) - _energy_at(curve, w0)) / (w1 * w1 - w0 * w0)
    mean = _energy_at(curve, max_crush) / max_crush if max_crush > 0.0 else 0.0
    v_end = float(v[-1])
    return BarrierMetrics(kw, max_crush, peak, mean, -v_end / v0, v_end - v0, curve)


def kw_window(v: _abi.Vehicle, window=(0.025, 0.4)) -> float:
    """CrushCurve::kw_window (src/vehicle.rs:60-64) of the bilinear curve."""
    w0, w1 = window
    return v.curve[1] + 2.0 * (w1 - w0) / (w1 * w1 - w0 * w0) * v.curve[0]


def barrier_prediction(v: _abi.Vehicle, speed: float) -> tuple[float, float]:
    """Vehicle::barrier_prediction (src/vehicle.rs:518-521): the crush that
    absorbs 1/2 m v^2 on the bilinear curve, and the force there."""
    a, b = 0.5 * v.curve[1], v.curve[0]
    e = 0.5 * v.mass * speed * speed
    c = (-b + math.sqrt(b * b + 4.0 * a * e)) / (2.0 * a)
    return c, v.curve[0] + v.curve[1] * c


@dataclass
class BarrierTargets:
    """BarrierTargets (src/vehicle.rs:983-1004)."""
    kw400: float
    max_crush: float
    restitution: float
    speed: float
    window: tuple = (0.025, 0.4)

    @classmethod
    def from_curve(cls, v: _abi.Vehicle, speed: float, window=(0.025, 0.4), restitution: float = 0.12):
        return cls(kw_window(v, window), barrier_prediction(v, speed)[0], restitution, speed, tuple(window))

    @classmethod
    def from_vehicle(cls, v: _abi.Vehicle, restitution: float = 0.12):
        return cls.from_curve(v, 35.0 * 0.44704, (0.025, 0.4), restitution)


@dataclass
class BarrierFit:
    """What calibrate returns: the tuned vehicle, the metrics of the last
    evaluated vehicle, and one (kw400, max_crush, restitution) per
    iteration."""
    vehicle: _abi.Vehicle
    metrics: BarrierMetrics
    iterations: int
    history: list = field(default_factory=list)


def barrier_models(jobs, element_size: float) -> list:
    """The rigid-barrier runs of calibrate (barrier_model with
    history_steps = 2, src/vehicle.rs:951-973)."""
    return [Model.barrier(v, element_size=element_size, speed=t.speed, end_time=BARRIER_END).set(history_steps=2)
            for v, t in jobs]


def calibrate_many(jobs: Sequence, element_size: float = TUNED_ELEMENT_SIZE, iterations: int = 8,
                   solve: Optional[Callable] = None, measure: Optional[Callable] = None,
                   log: Optional[Callable[[str], None]] = None) -> list:
    """calibrate (src/vehicle.rs:1006-1042) for several (vehicle, targets)
    pairs at once: every iteration solves all unfinished vehicles' barrier
    tests as one batch, then applies the reference's log-space secant step
    to each (stiffness x r_k^0.7, yield x (1/r_c)^0.7 r_k^0.21, and once
    KW400 and crush are within 15 %, modulus x clamp((1/r_e)^1.4, 0.5, 2)
    with the body modulus following); a vehicle stops when KW400 and crush
    are within 2 % and restitution within 10 %.

    `measure` (list of vehicles and targets -> list of BarrierMetrics)
    replaces the barrier solves altogether; `solve` (list of models ->
    results) replaces only the batched GPU solve."""
    if measure is None:
        run = solve or solve_batched

        def measure(pairs):
            models = barrier_models(pairs, element_size)
            return [barrier_metrics(mdl, r, v.name.decode(), 0, t.window)
                    for mdl, r, (v, t) in zip(models, run(models), pairs)]

    tuned = [copy_vehicle(v) for v, _ in jobs]
    last: list = [None] * len(jobs)
    hist: list = [[] for _ in jobs]
    active = list(range(len(jobs)))
    done_at = [0] * len(jobs)
    for it in range(iterations):
        if not active:
            break
        ms = measure([(tuned[i], jobs[i][1]) for i in active])
        still = []
        for i, m in zip(active, ms):
            t = jobs[i][1]
            last[i] = m
            done_at[i] = it + 1
            hist[i].append((m.kw400, m.max_crush, m.restitution))
            if log:
                log(f"calibrate {tuned[i].name.decode()} it {it}: KW {m.kw400 / 1e3:.0f}/{t.kw400 / 1e3:.0f} N/mm, "
                    f"crush {m.max_crush * 1e3:.0f}/{t.max_crush * 1e3:.0f} mm, e {m.restitution:.3f}/{t.restitution:.3f}")
            r_c = t.max_crush / m.max_crush
            r_k = t.kw400 / m.kw400 if m.max_crush >= t.window[1] else 1.0 / max(r_c, 1.0)
            r_e = t.restitution / max(m.restitution, 1e-3)
            if abs(r_k - 1.0) < 0.02 and abs(r_c - 1.0) < 0.02 and abs(r_e - 1.0) < 0.1:
                continue
            damp = 0.7
            v = tuned[i]
            v.curve[1] *= r_k ** damp
            v.curve[0] *= (1.0 / r_c) ** damp * r_k ** (0.3 * damp)
            if abs(r_k - 1.0) < 0.15 and abs(r_c - 1.0) < 0.15:
                v.modulus *= min(max((1.0 / r_e) ** (2.0 * damp), 0.5), 2.0)
                v.body_modulus = v.modulus
            still.append(i)
        active = still
    return [BarrierFit(tuned[i], last[i], done_at[i], hist[i]) for i in range(len(jobs))]


def calibrate(v: _abi.Vehicle, targets: BarrierTargets, element_size: float = TUNED_ELEMENT_SIZE,
              iterations: int = 8, solve: Optional[Callable] = None, measure: Optional[Callable] = None,
              log: Optional[Callable[[str], None]] = None) -> BarrierFit:
    """calibrate (src/vehicle.rs:1006-1042) for one vehicle."""
    return calibrate_many([(v, targets)], element_size, iterations, solve, measure, log)[0]
