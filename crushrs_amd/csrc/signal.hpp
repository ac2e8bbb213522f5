// Crash-pulse processing on the host (src/signal.rs counterparts); the C
// ABI is include/crushrs_signal.h.
#pragma once

#include <cstddef>
#include <vector>

#include "../../include/crushrs.h"

namespace crush {

// SAE J211 CFC filter, forward and backward (src/signal.rs:9-39).
std::vector<double> cfc_filter(const std::vector<double>& x, double dt, double cfc);
// Linear resampling onto t0 + k dt up to t_end (src/signal.rs:43-64);
// writes min(count, cap) values and returns the count.
size_t resample(const double* t, const double* y, size_t n, double t0, double dt, double t_end, double* out,
                size_t cap);
// Cumulative trapezoid from y0 (src/signal.rs:67-76).
void integrate(const double* t, const double* y, size_t n, double y0, double* out);
// Velocity change over a pulse sampled at k dt, k = 0 .. n-1, after a CFC
// filter of class cfc (none when cfc <= 0): Pulse::filtered then
// Pulse::velocity(0).last() (src/signal.rs:184-193).
double pulse_velocity_change(const double* accel, size_t n, double dt, double cfc);

struct Prepared;

// Pulse outputs of a finished solve from the device pulses of its groups
// ([min(G, 8)][P.n_pulse] doubles at dev_pulse, `samples` filled): the
// CFC 60 pulse delta-v per group into out->groups (when given) and the raw
// pulses into out->pulse (when given).
// t_end is the time the solve reached.
void pulse_results(const Prepared& P, const double* dev_pulse, int samples, double t_end, crush_results* out);
// The same from a host copy of the pulses (pulse_download).
void pulse_results_host(const Prepared& P, const double* host_pulse, int samples, double t_end,
                        crush_results* out);
std::vector<double> pulse_download(const Prepared& P, const double* dev_pulse);

}  // namespace crush
