"""Crash pulses: the reference's `signal::Pulse` (src/signal.rs:80-213) and
`PulseComparison` (src/vehicle.rs:1047-1088) over the library's signal ABI
(include/crushrs_signal.h): CFC filtering, resampling and integration run
in libcrushrs_amd; this module holds the arrays and reads NHTSA records."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _abi
from ._abi import check, load

G = 9.81  # m/s^2 per g (src/signal.rs:116)
PULSE_DT = 1e-4  # src/vehicle.rs:1102
PULSE_END = 0.15  # src/vehicle.rs:1104


def _p(a: np.ndarray):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_abi.f64p)


def cfc_filter(x, dt: float, cfc: float) -> np.ndarray:
    """cfc_filter (src/signal.rs:9-39)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    check(load().crush_cfc_filter(_p(x), len(x), dt, cfc, out.ctypes.data_as(_abi.f64p)))
    return out


def resample(t, y, t0: float, dt: float, t_end: float) -> tuple[np.ndarray, np.ndarray]:
    """resample (src/signal.rs:43-64)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    cap = int(np.floor((t_end - t0) / dt)) + 2
    out = np.empty(cap)
    n = C.c_int32()
    check(load().crush_resample(_p(t), _p(y), len(t), t0, dt, t_end, out.ctypes.data_as(_abi.f64p), cap, C.byref(n)))
    tu = t0 + np.arange(n.value) * dt
    return tu, out[:n.value].copy()


def integrate(t, y, y0: float) -> np.ndarray:
    """integrate (src/signal.rs:67-76)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty(len(t))
    check(load().crush_integrate(_p(t), _p(y), len(t), y0, out.ctypes.data_as(_abi.f64p)))
    return out


@dataclass
class PulseMetrics:
    """PulseMetrics (src/signal.rs:216-240)."""
    t_max_crush: float
    max_crush: float
    v_end: float
    restitution: float
    peak_accel: float


@dataclass
class Pulse:
    """A uniformly sampled acceleration history (m/s^2), t = 0 at first
    contact (src/signal.rs:80-213)."""
    time: np.ndarray
    accel: np.ndarray

    def dt(self) -> float:
        n = len(self.time)
        return (self.time[-1] - self.time[0]) / (n - 1) if n > 1 else 0.0

    @classmethod
    def parse_tsv(cls, text: str, in_g: bool) -> "Pulse":
        """Pulse::parse_tsv: `time value` rows, t >= 0 kept (src/signal.rs:104-125)."""
        time, accel = [], []
        for i, line in enumerate(text.splitlines()):
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            parts = [p for p in line.replace(",", " ").split() if p]
            try:
                t, a = float(parts[0]), float(parts[1])
            except (IndexError, ValueError):
                raise ValueError(f"line {i + 1}: bad row '{line}'")
            if t < -1e-12:
                continue
            time.append(t)
            accel.append(a * G if in_g else a)
        if len(time) < 4:
            raise ValueError("signal has fewer than 4 samples at t >= 0")
        return cls(np.array(time), np.array(accel))

    @classmethod
    def read_nhtsa_tsv(cls, path, in_g: bool = True) -> "Pulse":
        """Pulse::read_nhtsa_tsv (src/signal.rs:98-102)."""
        return cls.parse_tsv(Path(path).read_text(), in_g)

    @classmethod
    def from_results(cls, results, group: int) -> "Pulse":
        """The device-recorded pulse of a group (the reference's
        Pulse::from_history at PULSE_DT, src/signal.rs:127-146), unfiltered,
        in the accelerometer's local x."""
        if results.pulse is None or results.pulse_dt <= 0.0:
            raise ValueError("the solve recorded no pulse (settings.pulse_dt = 0)")
        a = results.pulse[group]
        return cls(np.arange(len(a)) * results.pulse_dt, a.copy())

    def resampled(self, dt: float) -> "Pulse":
        t, y = resample(self.time, self.accel, self.time[0], dt, self.time[-1])
        return Pulse(t, y)

    def forward(self) -> "Pulse":
        """Pulse::forward (src/signal.rs:156-163)."""
        dv = integrate(self.time, self.accel, 0.0)[-1]
        return Pulse(self.time.copy(), -self.accel) if dv > 0.0 else Pulse(self.time.copy(), self.accel.copy())

    @staticmethod
    def average(pulses) -> "Pulse":
        """Pulse::average (src/signal.rs:165-176)."""
        base = pulses[0]
        acc = np.zeros(len(base.time))
        for p in pulses:
            _, y = resample(p.time, p.accel, base.time[0], base.dt(), base.time[-1])
            acc += y / len(pulses)
        return Pulse(base.time.copy(), acc)

    @classmethod
    def read_nhtsa_list(cls, paths, in_g: bool = True) -> "Pulse":
        """Pulse::read_nhtsa_list (src/signal.rs:178-182)."""
        return cls.average([cls.read_nhtsa_tsv(p, in_g).forward() for p in paths])

    def negated(self) -> "Pulse":
        return Pulse(self.time.copy(), -self.accel)

    def filtered(self, cfc: float) -> "Pulse":
        return Pulse(self.time.copy(), cfc_filter(self.accel, self.dt(), cfc))

    def velocity(self, v0: float) -> np.ndarray:
        return integrate(self.time, self.accel, v0)

    def displacement(self, v0: float) -> np.ndarray:
        return integrate(self.time, self.velocity(v0), 0.0)

    def truncated(self, t_end: float) -> "Pulse":
        n = next((i for i, t in enumerate(self.time) if not t <= t_end + 1e-12), len(self.time))
        return Pulse(self.time[:n].copy(), self.accel[:n].copy())

    def metrics(self, mass: float, v0: float) -> PulseMetrics:
        m = _abi.PulseMetrics()
        check(load().crush_pulse_metrics_get(_p(self.time), _p(self.accel), len(self.time), mass, v0, C.byref(m)))
        return PulseMetrics(m.t_max_crush, m.max_crush, m.v_end, m.restitution, m.peak_accel)


@dataclass
class PulseComparison:
    """PulseComparison::new (src/vehicle.rs:1059-1077): both pulses on the
    measured pulse's time grid."""
    time: np.ndarray
    a_meas: np.ndarray
    a_sim: np.ndarray
    v_meas: np.ndarray
    v_sim: np.ndarray
    x_meas: np.ndarray
    x_sim: np.ndarray
    meas: PulseMetrics
    sim: PulseMetrics

    @classmethod
    def new(cls, meas: Pulse, sim: Pulse, mass: float, speed: float) -> "PulseComparison":
        t_end = min(meas.time[-1], sim.time[-1])
        m = meas.truncated(t_end)
        s = sim.resampled(m.dt()).truncated(t_end) if abs(sim.dt() - m.dt()) > 1e-9 * m.dt() else sim.truncated(t_end)
        n = min(len(m.time), len(s.time))
        return cls(m.time[:n], m.accel[:n], s.accel[:n], m.velocity(speed)[:n], s.velocity(speed)[:n],
                   m.displacement(speed)[:n], s.displacement(speed)[:n], m.metrics(mass, speed), s.metrics(mass, speed))

    def rms_accel_error(self) -> float:
        return float(np.sqrt(np.mean((self.a_meas - self.a_sim) ** 2)))
