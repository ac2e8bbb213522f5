/*
 * crushrs_signal.h — crash-pulse processing of the MI355X crash-solver
 * core: SAE J211 channel-frequency-class filtering, resampling, integration
 * and the barrier-pulse summary, the host-side counterparts of
 * src/signal.rs of osullivryan/crushrs.  Used on the accelerometer pulses
 * the solver records on the device (crush_settings.pulse_dt) and on
 * measured pulses (e.g. NHTSA test records).
 *
 * All arrays are caller-owned; functions return CRUSH_OK or an error code
 * (crush_last_error() has the message).
 */
#ifndef CRUSHRS_SIGNAL_H
#define CRUSHRS_SIGNAL_H

#include <stdint.h>

#include "crushrs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cfc_filter (src/signal.rs:9-39): phaseless 2-pole Butterworth run forward
 * and backward, each pass padded by reflection about the ends; CFC 60 is
 * -3 dB at about 100 Hz.  Signals of fewer than 4 samples are copied.
 * out may alias x. */
int32_t crush_cfc_filter(const double* x, int32_t n, double dt, double cfc, double* out);

/* resample (src/signal.rs:43-64): linear resampling of (t, y) (t ascending)
 * onto t0 + k dt, k = 0 .. floor((t_end - t0) / dt); *n_out is that count,
 * at most cap values are written (CRUSH_E_INPUT if cap is smaller). */
int32_t crush_resample(const double* t, const double* y, int32_t n, double t0, double dt, double t_end,
                       double* out, int32_t cap, int32_t* n_out);

/* integrate (src/signal.rs:67-76): cumulative trapezoid from y0. */
int32_t crush_integrate(const double* t, const double* y, int32_t n, double y0, double* out);

/* PulseMetrics::new (src/signal.rs:230-240) of a uniformly sampled pulse
 * (time, accel in m/s^2, forward positive) of a body of mass at initial
 * speed v0. */
typedef struct crush_pulse_metrics {
    double t_max_crush;  /* time of maximum crush (s)                        */
    double max_crush;    /* maximum of the double integral (m)               */
    double v_end;        /* velocity at the end of the record (m/s)          */
    double restitution;  /* max(-min velocity / v0, 0)                       */
    double peak_accel;   /* most negative acceleration (m/s^2)               */
} crush_pulse_metrics;

int32_t crush_pulse_metrics_get(const double* time, const double* accel, int32_t n, double mass, double v0,
                                crush_pulse_metrics* out);

#ifdef __cplusplus
}
#endif
#endif
