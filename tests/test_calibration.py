"""Host logic of the calibration drivers (crushrs_amd.calibration), on the
CPU: pulse metrics against closed forms, the probe rules of
calibrate_pulse (src/vehicle.rs:1144-1258) and the differential evolution
of calibrate_rail_boxes (src/headon.rs:135-250), driven by analytic scores
in place of solves so that every branch of the search is checked without a
GPU."""
import math

import numpy as np
import pytest

import crushrs_amd as ca
import oracle_ffi as of
from crushrs_amd import calibration as cal
from crushrs_amd import signal as sg


def harmonic_pulse(k=4e5, m=1000.0, v0=15.0, dt=1e-4, t_end=0.15):
    """A mass on a linear spring, F = k x: x = v0/w sin(wt), a = -k/m x."""
    w = math.sqrt(k / m)
    t = np.arange(int(round(t_end / dt)) + 1) * dt
    return sg.Pulse(t, -(v0 * w) * np.sin(w * t)), w


def test_force_crush_and_mean_force_of_a_linear_spring():
    """F = k x sampled along the crush: the loading phase ends at the
    quarter period with crush v0/w; the trapezoid mean force over the
    samples in [x0, x1] is k (first + last)/2 of their crush, and the
    energy-equivalent stiffness k (p1^2 - p0^2)/(x1^2 - x0^2) for the
    extreme samples p0, p1 in the window."""
    k, m, v0 = 4e5, 1000.0, 15.0
    p, w = harmonic_pulse(k, m, v0)
    fc = cal.force_crush(p, m, v0)
    t_max = fc.shape[0] - 1
    assert abs(p.time[t_max] - 0.5 * math.pi / w) <= 1e-4
    assert abs(fc[-1, 0] - v0 / w) < 1e-5 * v0 / w
    assert np.all(np.diff(fc[:, 0]) > 0)
    for x0, x1 in [(0.0, 0.1), (0.05, 0.2), (0.1, 0.23)]:
        inside = fc[(fc[:, 0] >= x0) & (fc[:, 0] <= x1), 0]
        assert abs(cal.mean_force(fc, x0, x1) / (0.5 * k * (inside[0] + inside[-1])) - 1.0) < 1e-5
    inside = fc[(fc[:, 0] >= 0.025) & (fc[:, 0] <= 0.2), 0]
    want = k * (inside[-1] ** 2 - inside[0] ** 2) / (0.2 ** 2 - 0.025 ** 2)
    assert abs(cal.kw(fc, 0.025, 0.2) / want - 1.0) < 1e-5 and abs(want / k - 1.0) < 0.02
    # windows the record does not reach, or that hold fewer than two samples
    assert cal.kw(fc, 0.025, 0.85) is None  # max crush 0.75 m
    assert cal.mean_force(fc, 0.8, 0.9) is None
    assert cal.mean_force(fc, 0.1, 0.1) is None


def test_constant_deceleration_force_and_initial_table():
    """A constant deceleration A: force m A at every crush, crush v0^2/(2A);
    the initial table spreads N_KNOTS knots over 0.95 of it, each at m A
    (the first at 1 kN when its window is one sample)."""
    m, v0, a = 1200.0, 15.0, 200.0
    dt = 1e-4
    t = np.arange(int(round(0.15 / dt)) + 1) * dt
    p = sg.Pulse(t, np.where(t <= v0 / a, -a, 0.0))
    fc = cal.force_crush(p, m, v0)
    x_max = v0 * v0 / (2 * a)
    assert abs(fc[-1, 0] - x_max) < 1e-3
    assert abs(cal.mean_force(fc, 0.1, 0.4) - m * a) < 1e-6 * m * a
    table = cal.initial_force_table(p, m, v0)
    assert table.shape == (cal.N_KNOTS, 2)
    assert np.allclose(table[:, 0], np.arange(cal.N_KNOTS) * fc[-1, 0] * 0.95 / (cal.N_KNOTS - 1))
    assert np.allclose(table[1:, 1], m * a, rtol=1e-6)
    assert table[0, 1] >= 1e3


def test_initial_table_is_non_decreasing_from_the_second_knot():
    """A force that falls with crush keeps knot 0 and 1 as measured and
    raises the rest to the running maximum from knot 1 on."""
    m, v0 = 1000.0, 12.0
    dt = 1e-4
    t = np.arange(1501) * dt
    p = sg.Pulse(t, -300.0 * np.exp(-t / 0.01))
    table = cal.initial_force_table(p, m, v0)
    f = table[:, 1]
    assert np.all(np.diff(f[1:]) >= 0)
    assert f[0] > f[1]  # the first knot is not raised


def test_pulse_objective_closed_form():
    """Identical pulses score 0; a simulated pulse offset by a constant da
    has velocity error da*t: J = da*sqrt(mean(t^2)) + 0.5*da*T + 0.1*da/9.81."""
    p, _ = harmonic_pulse()
    m, v0 = 1000.0, 15.0
    assert cal.pulse_objective(sg.PulseComparison.new(p, p, m, v0)) == 0.0
    da = 3.0
    q = sg.Pulse(p.time.copy(), p.accel + da)
    cmp = sg.PulseComparison.new(p, q, m, v0)
    t = cmp.time
    want = da * math.sqrt(np.mean(t * t)) + 0.5 * da * t[-1] + 0.1 * da / 9.81
    assert abs(cal.pulse_objective(cmp) - want) < 1e-9 * want


def table_vehicle(forces, modulus=4e6, body=2.7e8):
    v = ca.vehicle("neon")
    cal.set_force_table(v, [[0.1 * i, f] for i, f in enumerate(forces)])
    v.modulus, v.body_modulus = modulus, body
    return v


def test_pulse_candidate_rules():
    v = table_vehicle([10e3, 100e3, 120e3, 150e3, 150e3, 300e3, 310e3])
    up = cal.pulse_candidate(v, 2, math.exp(0.3))
    f = cal.force_table(up)[:, 1]
    assert f[2] == 120e3 * math.exp(0.3) and f[3] == f[2]  # raised to non-decreasing
    assert np.all(np.diff(f[1:]) >= 0)
    # knot 0 may sit above knot 1: only from the second knot on is the table raised
    hi = cal.pulse_candidate(v, 0, 20.0)
    assert cal.force_table(hi)[0, 1] == 200e3 and cal.force_table(hi)[1, 1] == 100e3
    # a move the monotone rule undoes is no probe
    flat = table_vehicle([10e3, 100e3, 100e3, 100e3, 100e3, 100e3, 100e3])
    assert cal.pulse_candidate(flat, 3, math.exp(-0.3)) is None
    assert cal.pulse_candidate(flat, 3, math.exp(0.3)) is not None
    # moduli, clamped
    assert cal.pulse_candidate(v, cal.N_KNOTS, 1e3).modulus == cal.MODULUS_RANGE[1]
    assert cal.pulse_candidate(v, cal.N_KNOTS, 1e-3).modulus == cal.MODULUS_RANGE[0]
    assert cal.pulse_candidate(v, cal.N_KNOTS + 1, 1e3).body_modulus == cal.BODY_MODULUS_RANGE[1]
    assert cal.pulse_candidate(v, cal.N_KNOTS + 1, 1.0 / 1.5).body_modulus == 2.7e8 / 1.5
    # the base is untouched
    assert cal.force_table(v)[2, 1] == 120e3 and v.modulus == 4e6


def log_params(v):
    return np.log(np.concatenate([cal.force_table(v)[:, 1], [v.modulus, v.body_modulus]]))


def quadratic_score(target, calls):
    def score(vehicles):
        calls.append(len(vehicles))
        return [(float(np.sum((log_params(v) - target) ** 2)), None) for v in vehicles]
    return score


def test_pattern_search_on_a_quadratic_bowl():
    """calibrate_pulse's search with an analytic score (squared distance in
    log space to a point the moves can reach): the first batch holds the
    start and 18 probes, later ones 18; the objective never rises; the step
    halves only on a round without improvement and the search stops below
    0.04; it ends on the target grid point."""
    base = table_vehicle([20e3, 50e3, 100e3, 200e3, 400e3, 800e3, 1600e3])
    start = log_params(base)
    moves = np.array([3, 1, -1, 1, -1, 1, 0, -2, 2]) * 0.3
    target = start + moves
    assert np.all(np.diff(np.exp(target[1:7])) >= 0)  # reachable through monotone tables
    calls = []
    fit = cal.calibrate_pulse(base, harmonic_pulse()[0], 15.0, rounds=200, score=quadratic_score(target, calls))
    assert calls[0] == 19 and all(c <= 18 for c in calls[1:])
    assert fit.batches == len(calls) == fit.rounds
    assert fit.evaluations == sum(calls)
    js = [h[1] for h in fit.history]
    assert all(b <= a for a, b in zip(js, js[1:]))
    steps = [0.3] + [h[2] for h in fit.history]
    for (rnd, j, s, acc), s_prev in zip(fit.history, steps):
        assert s == (s_prev if acc is not None else 0.5 * s_prev)
    assert fit.history[-1][3] is None and fit.history[-1][2] < 0.04
    assert fit.objective < 1e-18
    assert np.allclose(log_params(fit.vehicle), target, atol=1e-9)


def test_pattern_search_takes_the_first_best_probe_and_needs_strict_improvement():
    """Probes are ordered knot 0 up, knot 0 down, knot 1 up, ..., body
    modulus down; of equal best probes the first wins, and a probe no better
    than the current point is not taken."""
    base = table_vehicle([20e3, 100e3, 120e3, 150e3, 200e3, 300e3, 310e3])
    seen = []

    def score(vehicles):
        seen.append([log_params(v) for v in vehicles])
        if len(seen) == 1:  # start + probes: every probe equally better
            return [(1.0, None)] + [(0.5, None)] * (len(vehicles) - 1)
        return [(0.5, None)] * len(vehicles)  # nothing strictly better

    fit = cal.calibrate_pulse(base, harmonic_pulse()[0], 15.0, rounds=2, score=score)
    assert fit.history[0][3] == (0, 1.0)  # knot 0 up
    assert fit.history[1][3] is None and fit.history[1][2] == 0.15
    assert abs(fit.vehicle.force_table[0][1] - 20e3 * math.exp(0.3)) < 1e-6


def test_calibrate_pulse_initialises_the_table_from_the_measured_pulse():
    v = ca.vehicle("neon")
    v.n_force_table = 0
    v.n_compaction_map = 3
    p, _ = harmonic_pulse()
    got = []

    def score(vehicles):
        got.extend(vehicles)
        return [(1.0, None)] * len(vehicles)

    fit = cal.calibrate_pulse(v, p, 15.0, rounds=1, score=score)
    meas = p.filtered(60.0).resampled(sg.PULSE_DT).truncated(sg.PULSE_END)
    assert np.array_equal(cal.force_table(got[0]), cal.initial_force_table(meas, v.mass, 15.0))
    assert got[0].n_compaction_map == 0 and fit.vehicle.n_compaction_map == 0


def test_xorshift64star_sequence_properties():
    """xorshift64* (Vigna 2016): the state map is the xorshift 12/25/27 step
    of a nonzero 64-bit state, the output its product with
    0x2545F4914F6CDD1D; uniform() lies in [0, 1) with mean 1/2 and
    below(n) covers 0..n-1 evenly."""
    r = cal.XorShift64Star(0x9E3779B97F4A7C15)
    x = 0x9E3779B97F4A7C15
    for _ in range(100):
        x ^= x >> 12
        x ^= (x << 25) % (1 << 64)
        x ^= x >> 27
        assert r.next_u64() == (x * 0x2545F4914F6CDD1D) % (1 << 64)
    u = np.array([r.uniform() for _ in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    counts = np.bincount([r.below(7) for _ in range(14000)], minlength=7)
    assert counts.min() > 1800 and counts.max() < 2200
    # the state never returns to zero and the first 10^4 states are distinct
    s = cal.XorShift64Star(1)
    states = set()
    for _ in range(10000):
        s.next_u64()
        assert s.state != 0
        states.add(s.state)
    assert len(states) == 10000


def test_rail_box_unit_cube_round_trip():
    v = ca.vehicle("navigator")
    for u in ([0.5, 0.5, 0.5, 0.5], [0.1, 0.9, 0.2, 0.0], [1.0, 0.0, 0.0, 1.0]):
        w = cal.with_rail_box(v, u)
        assert np.allclose(cal.unit_from_rail_box(w), u, atol=1e-12)
    width, (z0, z1), frac = cal.rail_box_from_unit(v, [0.0, 1.0, 1.0, 1.0])
    assert width == 0.25 * v.size[1] and frac == 0.95 and z0 == 0.5 * v.size[2] - 0.3 * v.size[2] and z1 <= v.size[2]
    bare = ca.vehicle("neon")
    assert not bare.has_rail_box
    d = cal.with_default_rail_box(bare)
    assert d.has_rail_box and d.rail_box_width == 0.5 * bare.size[1] and d.rail_box_force_fraction == 0.7
    assert np.allclose(cal.unit_from_rail_box(d), [(0.5 - 0.25) / 0.7, (0.4 - 0.15) / 0.35, (0.4 - 0.15) / 0.45,
                                                   0.2 / 0.45])


def test_differential_evolution_on_a_quadratic_bowl():
    """calibrate_rail_boxes' search with an analytic score: the seed is the
    vehicles' boxes, each generation is one batch of `population` trials in
    the unit cube, a member is replaced only by a trial no worse, so the
    best never rises; runs are repeatable; the fit's vehicles carry the
    best member's boxes."""
    a, b = ca.vehicle("navigator"), ca.vehicle("neon-pulse")
    target = np.array([0.3, 0.6, 0.4, 0.8, 0.7, 0.2, 0.5, 0.35])
    calls = []

    def score(us):
        calls.append([list(u) for u in us])
        return [(float(np.sum((np.array(u) - target) ** 2)), u, None) for u in us]

    setup = cal.HeadOnSetup(13.5, 13.5)
    fit = cal.calibrate_rail_boxes(a, b, setup, None, None, population=12, generations=40, score=score)
    assert calls[0][0] == cal.unit_from_rail_box(a) + cal.unit_from_rail_box(b)
    assert len(calls) == 41 and all(len(c) == 12 for c in calls)
    assert all(0.0 <= x <= 1.0 for c in calls for u in c for x in u)
    best = [h[0] for h in fit.history]
    assert all(y <= x for x, y in zip(best, best[1:]))
    assert fit.objective < 0.05 * best[0]
    assert fit.evaluations == 12 * 41 and fit.batches == 41
    u = np.array(fit.cmp_a)
    assert np.allclose(cal.unit_from_rail_box(fit.a) + cal.unit_from_rail_box(fit.b), u, atol=1e-12)
    calls2 = []

    def score2(us):
        calls2.append([list(x) for x in us])
        return score(us)

    fit2 = cal.calibrate_rail_boxes(a, b, setup, None, None, population=12, generations=40, score=score2)
    assert calls2 == calls[41:] and fit2.objective == fit.objective


def test_initial_table_reproduces_the_reference_pulse_fit_of_nhtsa_2320():
    """The neon-pulse preset's table is the reference's calibrate_pulse
    result on NHTSA test 2320 at 35 mph (src/vehicle.rs:529-560; knots to
    0.1 mm, forces to 1 kN).  Its knots are where initial_force_table puts
    them (0.95 of the measured crush over six spacings), and the forces of
    knots 1, 3, 4, 5 and 6 are the initial mean forces: the search moved
    only knot 0 (up) and knot 2 (down, then raised back to knot 1 by the
    non-decreasing rule)."""
    from pathlib import Path
    golden = Path(__file__).resolve().parent / "golden" / "nhtsa" / "v02320tsv.078"
    preset = ca.vehicle("neon-pulse")
    meas = sg.Pulse.read_nhtsa_tsv(golden, True).filtered(60.0).resampled(sg.PULSE_DT).truncated(sg.PULSE_END)
    table = cal.initial_force_table(meas, preset.mass, 35.0 * ca.MPH)
    want = cal.force_table(preset)
    # the preset's spacing is 0.11652 m, this record gives 0.116548 m (a
    # 0.2 mm difference in the 736 mm maximum crush, as from a speed 0.002
    # m/s lower); a dropped 0.95, another knot count or another spacing
    # moves the last knot by centimetres
    assert np.abs(table[:, 0] - want[:, 0]).max() < 0.3e-3
    for i in (1, 3, 4, 5, 6):
        assert abs(table[i, 1] / want[i, 1] - 1.0) < 5e-3, (i, table[i, 1], want[i, 1])
    assert want[0, 1] > table[0, 1] and want[2, 1] == want[1, 1] < table[2, 1]


# ---------------------------------------------------------------- barrier metrics and calibrate

class _Parts:
    """A stand-in model for barrier_metrics: part names only."""

    def __init__(self, names):
        self._names = names

    def part_names(self):
        return list(self._names)


class _History:
    def __init__(self, t, parts):
        self.history_time = t
        self.parts = parts


def spring_history(k=625e3, m=1000.0, v0=15.0, dt=1e-4, t_end=0.2, split=0.3):
    """A body moving +x at v0 into a linear spring F = k x from t = 0, as a
    two-part vehicle history (masses split and 1 - split, same motion)."""
    w = math.sqrt(k / m)
    t = np.arange(int(round(t_end / dt)) + 1) * dt
    tc = math.pi / w
    x = np.where(t <= tc, v0 / w * np.sin(w * np.minimum(t, tc)), -v0 * (t - tc))
    v = np.where(t <= tc, v0 * np.cos(w * np.minimum(t, tc)), -v0)
    P = np.zeros((len(t), 3, 11))
    for p, mp in ((0, split * m), (1, (1 - split) * m), (2, 50.0)):
        P[:, p, 0] = mp
        if p < 2:
            P[:, p, 1], P[:, p, 4] = x, v
    return _History(t, P), w


def test_barrier_metrics_of_a_linear_spring():
    """barrier_metrics on a body into a linear spring k: contact where the
    force first exceeds 2 % of the peak (crush x_c there), loading to the
    maximum v0/w; KW over (w0, w1) is k (1 + 2 x_c (w1 - w0)/(w1^2 - w0^2))
    (the curve measured from x_c), peak force k v0/w, mean force
    k (x_max + x_c)/2 over the loading crush, restitution 1 and delta-v
    -2 v0.  Only the vehicle's own parts (<name>, <name>_crush) count."""
    k, m, v0 = 625e3, 1000.0, 15.0
    h, w = spring_history(k, m, v0)
    model = _Parts(["car", "car_crush", "wall"])
    bm = cal.barrier_metrics(model, h, "car")
    x = h.parts[:, 0, 1]
    f_exact = k * x
    xc = x[np.nonzero(f_exact > 0.02 * k * v0 / w)[0][0]]
    x_max = v0 / w
    assert abs(bm.max_crush - (x_max - xc)) < 2e-3
    assert abs(bm.peak_force / (k * x_max) - 1.0) < 1e-3
    w0, w1 = 0.025, 0.4
    want_kw = k * (1.0 + 2.0 * xc * (w1 - w0) / (w1 * w1 - w0 * w0))
    assert abs(bm.kw400 / want_kw - 1.0) < 2e-3, (bm.kw400, want_kw)
    assert abs(bm.mean_force / (0.5 * k * (x_max + xc)) - 1.0) < 2e-3
    assert abs(bm.restitution - 1.0) < 1e-9 and abs(bm.delta_v + 2.0 * v0) < 1e-9
    assert np.all(np.diff(bm.curve[:, 0]) > 0)
    with pytest.raises(ValueError):
        cal.barrier_metrics(model, h, "truck")


def test_bilinear_curve_targets():
    """CrushCurve::kw_window and Vehicle::barrier_prediction: KW of
    F = F_y + k x over (w0, w1) is k + 2 F_y (w1 - w0)/(w1^2 - w0^2); the
    predicted crush absorbs 1/2 m v^2 exactly."""
    v = ca.vehicle("neon-raw")
    fy, kk = v.curve[0], v.curve[1]
    assert abs(cal.kw_window(v) - (kk + 2 * fy * 0.375 / (0.16 - 0.025 ** 2))) < 1e-6 * kk
    c, f = cal.barrier_prediction(v, 15.6464)
    assert abs(fy * c + 0.5 * kk * c * c - 0.5 * v.mass * 15.6464 ** 2) < 1e-6 * v.mass
    assert abs(f - (fy + kk * c)) < 1e-9 * f
    t = cal.BarrierTargets.from_vehicle(v, 0.12)
    assert t.window == (0.025, 0.4) and abs(t.speed - 35 * 0.44704) < 1e-12 and t.max_crush == c


def test_calibrate_secant_converges_on_an_analytic_barrier():
    """calibrate_many with an analytic barrier (the vehicle behaves as its
    bilinear curve made 30 % stiffer: KW and crush of that curve;
    restitution falling with the modulus): both vehicles converge to their
    targets within the reference's tolerances; each iteration measures only
    the unfinished ones, in one call; the modulus moves only once KW and
    crush are within 15 %."""
    def fake(v, t):
        s = cal.copy_vehicle(v)
        s.curve[0] *= 1.3
        s.curve[1] *= 1.3
        kw = cal.kw_window(s, t.window)
        crush = cal.barrier_prediction(s, t.speed)[0]
        e = 0.12 * (v.modulus / 20e6) ** -0.5
        return cal.BarrierMetrics(kw, crush, 0.0, 0.0, e, 0.0, np.zeros((0, 2)))

    calls = []

    def measure(pairs):
        calls.append(len(pairs))
        return [fake(v, t) for v, t in pairs]

    jobs = []
    for name in ("neon-raw", "silverado-raw"):
        v = ca.vehicle(name)
        v.modulus = v.body_modulus = 10e6
        jobs.append((v, cal.BarrierTargets.from_vehicle(ca.vehicle(name), 0.12)))
    fits = cal.calibrate_many(jobs, iterations=30, measure=measure)
    assert calls[0] == 2 and all(c <= 2 for c in calls)
    for fit, (v0, t) in zip(fits, jobs):
        m = fit.metrics
        assert abs(m.kw400 / t.kw400 - 1) < 0.02 and abs(m.max_crush / t.max_crush - 1) < 0.02
        assert abs(m.restitution / t.restitution - 1) < 0.1
        assert fit.iterations < 30
        assert fit.vehicle.body_modulus == fit.vehicle.modulus
        es = [h[2] for h in fit.history]
        first_close = next(i for i, (kw, c, e) in enumerate(fit.history)
                           if abs(t.kw400 / kw - 1) < 0.15 and abs(t.max_crush / c - 1) < 0.15)
        assert first_close >= 1 and es[first_close] == es[0]  # the curve moves first, the modulus after
    assert jobs[0][0].modulus == 10e6  # inputs untouched


@pytest.mark.parametrize("preset,kw,crush,e", [("silverado", 2563e3, 0.540, 0.119)])
def test_tuned_silverado_barrier_metrics_match_the_reference(preset, kw, crush, e):
    """The reference's published metrics of its tuned Silverado at 0.3 m
    elements, 35 mph rigid barrier (src/vehicle.rs:635-637: KW400 2563
    N/mm, max crush 540 mm, restitution 0.119, against targets 2550 N/mm
    and 513 mm), from the reference's default lane path restated (the
    oracle's AVX2 path) and barrier_metrics.  The targets match exactly;
    the restated solve gives 2612 N/mm, 538.7 mm and 0.126 (its f64 path
    2614 N/mm, 538.5 mm, 0.123): 1.9 % above the published KW400, which
    the bars below allow (the published figures date from the preset's
    tuning and are not regenerated by the reference's tests)."""
    m = cal.barrier_models([(ca.vehicle(preset), cal.BarrierTargets(0, 0, 0, 35 * ca.MPH))], ca.TUNED_ELEMENT_SIZE)[0]
    r = of.run_lanes(m, threads=8)
    bm = cal.barrier_metrics(m, r, preset)
    assert abs(bm.kw400 / kw - 1.0) < 0.025, bm.kw400
    assert abs(bm.max_crush - crush) < 0.005, bm.max_crush
    assert abs(bm.restitution - e) < 0.01, bm.restitution
    t = cal.BarrierTargets.from_vehicle(ca.vehicle(preset + "-raw"), 0.12)
    assert round(t.kw400 / 1e3) == 2550 and round(t.max_crush * 1e3) == 513


def test_tuned_neon_reproduces_ncap_barrier_test():
    """tests/impact.rs:50-60 on the oracle's f64 path: the tuned Neon's
    simulated KW400 within 5 % and crush within 10 % of the NCAP Neon's
    curve, restitution between 0.03 and 0.2."""
    target = ca.vehicle("neon-raw")
    t = cal.BarrierTargets.from_vehicle(target, 0.12)
    m = cal.barrier_models([(ca.vehicle("neon"), t)], ca.TUNED_ELEMENT_SIZE)[0]
    r = of.run(m)
    bm = cal.barrier_metrics(m, r, "neon")
    assert abs(bm.kw400 - t.kw400) < 0.05 * t.kw400, (bm.kw400, t.kw400)
    assert abs(bm.max_crush - t.max_crush) < 0.10 * t.max_crush, (bm.max_crush, t.max_crush)
    assert 0.03 < bm.restitution < 0.2
