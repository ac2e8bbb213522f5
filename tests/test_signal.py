"""Crash-pulse processing (src/signal.rs) and the accelerometer pulse of a
solve, on the CPU.

The library's signal ABI (include/crushrs_signal.h) and the oracle's C
restatement are pinned by the reference's own signal tests
(src/signal.rs:306-345) and agree with each other; the oracle's recorded
pulse is pinned by the reference's NHTSA validations: test 2320 (rigid
barrier, tests/impact.rs:192-211 and the simulated-vs-measured figures of
src/vehicle.rs:533-538) and test 4429 (head-on pulse delta-v,
tests/impact.rs:284-309), against the measured records the reference's
tests read (tests/golden/nhtsa/)."""
from pathlib import Path

import numpy as np
import pytest

import crushrs_amd as ca
from crushrs_amd import signal as sg
import oracle_ffi as of

GOLDEN = Path(__file__).resolve().parent / "golden" / "nhtsa"


def oracle_filter(x, dt, cfc):
    x = np.ascontiguousarray(x, float)
    out = np.empty_like(x)
    of.load().oracle_cfc_filter(of.ptr(x), len(x), dt, cfc, of.ptr(out))
    return out


def oracle_resample(t, y, t0, dt, t_end):
    t = np.ascontiguousarray(t, float)
    y = np.ascontiguousarray(y, float)
    out = np.empty(int(np.floor((t_end - t0) / dt)) + 2)
    n = of.load().oracle_resample(of.ptr(t), of.ptr(y), len(t), t0, dt, t_end, of.ptr(out), len(out))
    return out[:n]


@pytest.mark.parametrize("impl", ["library", "oracle"])
def test_cfc60_passes_dc_and_kills_1khz(impl):
    # src/signal.rs:310-324
    f = sg.cfc_filter if impl == "library" else oracle_filter
    dt = 1e-4
    t = np.arange(5000) * dt
    y = f(10.0 + 5.0 * np.sin(2.0 * np.pi * 1000.0 * t), dt, 60.0)
    assert np.abs(y[1000:4000] - 10.0).max() < 0.05
    y100 = f(np.sin(2.0 * np.pi * 100.0 * t), dt, 60.0)
    assert abs(np.abs(y100[1000:4000]).max() - 0.707) < 0.05


def test_parse_and_integrate_tsv():
    # src/signal.rs:326-336
    text = "-0.001\t-0.1\r\n0\t0\n0.001\t-1\n0.002\t-1\n0.003\t-1\n0.004\t-1\n"
    p = sg.Pulse.parse_tsv(text, True)
    assert len(p.time) == 5
    assert abs(p.dt() - 0.001) < 1e-12
    v = p.velocity(10.0)
    assert abs(v[4] - (10.0 - 9.81 * 0.0035)) < 1e-9
    m = p.metrics(1000.0, 10.0)
    assert 0.039 < m.max_crush < 0.04


@pytest.mark.parametrize("impl", ["library", "oracle"])
def test_resample_is_linear(impl):
    # src/signal.rs:338-345
    t, y = [0.0, 1.0, 3.0], [0.0, 2.0, 6.0]
    yu = sg.resample(t, y, 0.0, 0.5, 3.0)[1] if impl == "library" else oracle_resample(t, y, 0.0, 0.5, 3.0)
    assert len(yu) == 7
    assert abs(yu[1] - 1.0) < 1e-12 and abs(yu[3] - 3.0) < 1e-12 and abs(yu[6] - 6.0) < 1e-12


def test_library_filter_and_resample_equal_the_oracle():
    """The library's signal functions against the oracle's restatement on a
    measured record and on noise: identical to round-off."""
    p = sg.Pulse.read_nhtsa_tsv(GOLDEN / "v02320tsv.078", True)
    r = p.resampled(sg.PULSE_DT)
    assert np.allclose(r.accel, oracle_resample(p.time, p.accel, p.time[0], sg.PULSE_DT, p.time[-1]), rtol=0,
                       atol=1e-12 * np.abs(p.accel).max())
    rng = np.random.default_rng(4)
    for x, dt in [(r.accel, sg.PULSE_DT), (rng.normal(size=3001), 8e-5), (rng.normal(size=7), 1e-3)]:
        a, b = sg.cfc_filter(x, dt, 60.0), oracle_filter(x, dt, 60.0)
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1.0)
    t = np.arange(200) * 1e-3
    assert np.allclose(sg.integrate(t, np.cos(t), 1.0), 1.0 + np.sin(t), atol=1e-6)


def test_measured_4429_records_average_forward():
    """Pulse::read_nhtsa_list: the two sill records of each vehicle in test
    4429, each turned forward-positive and averaged, decelerate the vehicle
    by about its NHTSA delta-v (9.9 / 19.6 m/s) over 150 ms."""
    nav = sg.Pulse.read_nhtsa_list([GOLDEN / "v04429tsv.179", GOLDEN / "v04429tsv.182"])
    neon = sg.Pulse.read_nhtsa_list([GOLDEN / "v04429tsv.089", GOLDEN / "v04429tsv.092"])
    for p, want in [(nav, 9.89), (neon, 19.57)]:
        f = p.filtered(60.0).resampled(sg.PULSE_DT).truncated(sg.PULSE_END)
        dv = -f.velocity(0.0)[-1]
        assert abs(dv - want) < 1.5, (dv, want)


def test_oracle_pulse_matches_nhtsa_2320_barrier():
    """tests/impact.rs:192-211 on the oracle: the pulse-calibrated Neon in
    a 35 mph rigid-barrier test against the measured rear-seat pulse of
    NHTSA test 2320 (CFC 60): delta-v within 5 %, peak within 20 %, crush
    within 15 %, RMS error under 8 g, velocity within 1.5 m/s throughout;
    and the reference's own simulated figures (src/vehicle.rs:536-538:
    delta-v 17.64 m/s, peak -32.9 g, crush 800 mm, RMS 6.2 g)."""
    v = ca.vehicle("neon-pulse")
    speed = 35.0 * ca.MPH
    m = ca.Model.barrier(v, element_size=ca.TUNED_ELEMENT_SIZE, speed=speed, end_time=sg.PULSE_END)
    r = of.run(m)
    sim = sg.Pulse.from_results(r, 0).filtered(60.0)
    meas = sg.Pulse.read_nhtsa_tsv(GOLDEN / "v02320tsv.078", True).filtered(60.0).resampled(
        sg.PULSE_DT).truncated(sg.PULSE_END)
    cmp = sg.PulseComparison.new(meas, sim, v.mass, speed)
    dv_meas, dv_sim = speed - cmp.v_meas[-1], speed - cmp.v_sim[-1]
    assert abs(dv_sim - dv_meas) < 0.05 * dv_meas
    assert abs(cmp.sim.peak_accel / cmp.meas.peak_accel - 1.0) < 0.2
    assert abs(cmp.sim.max_crush / cmp.meas.max_crush - 1.0) < 0.15
    assert cmp.rms_accel_error() < 8.0 * 9.81
    assert np.abs(cmp.v_meas - cmp.v_sim).max() < 1.5
    # the reference's published simulation of the same test
    assert abs(dv_sim - 17.64) < 0.1
    assert abs(cmp.sim.peak_accel / 9.81 - (-32.9)) < 2.0
    assert abs(cmp.sim.max_crush - 0.800) < 0.05
    assert abs(cmp.rms_accel_error() / 9.81 - 6.2) < 0.5
    # the library's pulse delta-v is the same quantity (speed - end velocity)
    assert abs(r.groups[0].pulse_delta_v - (speed - sim.velocity(speed)[-1])) < 1e-9


def test_oracle_pulse_delta_v_matches_nhtsa_4429():
    """tests/impact.rs:284-309 with the reference's delta-v: the CFC 60 rear
    accelerometer pulses of Navigator vs pulse-tuned Neon (1378 kg) at
    30.14 mph, integrated over 150 ms: 9.89 +- 1.0 and 19.57 +- 2.0 m/s."""
    a = ca.renamed(ca.vehicle("navigator"), "navigator_a")
    b = ca.vehicle("neon-pulse")
    b.mass = 1378.0
    b = ca.renamed(b, "neon_b")
    v0 = 30.14 * ca.MPH
    r = of.run(ca.Model.headon(a, b, v0, v0))
    dv_a, dv_b = r.groups[0].pulse_delta_v, r.groups[1].pulse_delta_v
    assert abs(dv_a - 9.89) < 1.0, dv_a
    assert abs(dv_b - 19.57) < 2.0, dv_b
    # the same from the recorded pulses, as HeadOn::run forms them
    pa = sg.Pulse.from_results(r, 0).filtered(60.0)
    pb = sg.Pulse.from_results(r, 1).negated().filtered(60.0)
    assert abs((v0 - pa.velocity(v0)[-1]) - dv_a) < 1e-9
    assert abs((v0 - pb.velocity(v0)[-1]) - dv_b) < 1e-9


def test_pulse_grid_covers_the_solve():
    """One sample per pulse_dt from t = 0 to the last step (the resampling
    of Pulse::from_history), and none when pulse_dt is 0."""
    m = ca.Model.crash(30.0, 30.0).set(end_time=0.02)
    r = of.run(m)
    assert r.pulse.shape[1] == int(np.floor(r.time / m.settings.pulse_dt)) + 1
    assert r.pulse[0][0] == 0.0 or abs(r.pulse[0][0]) < 1e-6
    r0 = of.run(ca.Model.crash(30.0, 30.0).set(end_time=0.02, pulse_dt=0.0))
    assert r0.pulse is None and all(g.pulse_delta_v == 0.0 for g in r0.groups)
