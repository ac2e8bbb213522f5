// Crash-pulse processing on the host (include/crushrs_signal.h): the
// counterparts of src/signal.rs, applied to the accelerometer pulses the
// device records and to measured pulses.
#include "signal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/crushrs_signal.h"
#include "common.hpp"
#include "hip_util.hpp"
#include "prep.hpp"

namespace crush {

// One causal pass of the 2-pole low-pass over x, started on a copy
// extended at both ends by reflection about the end values
// (src/signal.rs:23-35).
static std::vector<double> cfc_pass(const std::vector<double>& x, double dt, double cfc, const double* a,
                                    const double* b) {
    const size_t n = x.size();
    size_t pad = (size_t)std::ceil(3.0 / (cfc * dt));
    pad = std::max<size_t>(std::min(pad, n - 1), 1);
    std::vector<double> ext;
    ext.reserve(n + 2 * pad);
    for (size_t i = 0; i < pad; ++i) ext.push_back(2.0 * x[0] - x[pad - i]);
    ext.insert(ext.end(), x.begin(), x.end());
    for (size_t i = 1; i <= pad; ++i) ext.push_back(2.0 * x[n - 1] - x[n - 1 - i]);
    std::vector<double> y(ext.size(), 0.0);
    y[0] = ext[0];
    y[1] = ext[1];
    for (size_t i = 2; i < ext.size(); ++i)
        y[i] = a[0] * ext[i] + a[1] * ext[i - 1] + a[2] * ext[i - 2] + b[0] * y[i - 1] + b[1] * y[i - 2];
    return std::vector<double>(y.begin() + (long)pad, y.begin() + (long)(pad + n));
}

std::vector<double> cfc_filter(const std::vector<double>& x, double dt, double cfc) {
#pragma clang fp contract(off)
    if (x.size() < 4) return x;
    // SAE J211-1 digital CFC filter: design frequency 2.0775 x CFC,
    // prewarped bilinear 2-pole Butterworth (src/signal.rs:13-21)
    const double pi = 3.141592653589793;
    const double wd = 2.0 * pi * cfc * 2.0775;
    const double wa = std::sin(wd * dt / 2.0) / std::cos(wd * dt / 2.0);
    const double s2 = std::sqrt(2.0);
    const double den = 1.0 + s2 * wa + wa * wa;
    const double a0 = wa * wa / den;
    const double a[3] = {a0, 2.0 * a0, a0};
    const double b[2] = {-2.0 * (wa * wa - 1.0) / den, (-1.0 + s2 * wa - wa * wa) / den};
    std::vector<double> fwd = cfc_pass(x, dt, cfc, a, b);
    std::reverse(fwd.begin(), fwd.end());
    std::vector<double> back = cfc_pass(fwd, dt, cfc, a, b);
    std::reverse(back.begin(), back.end());
    return back;
}

size_t resample(const double* t, const double* y, size_t n, double t0, double dt, double t_end, double* out,
                size_t cap) {
#pragma clang fp contract(off)
    const size_t m = (size_t)std::floor((t_end - t0) / dt) + 1;
    size_t j = 0;
    for (size_t k = 0; k < m && k < cap; ++k) {
        const double tk = t0 + (double)k * dt;
        while (j + 1 < n && t[j + 1] < tk) ++j;
        double v;
        if (j + 1 >= n)
            v = y[n - 1];
        else if (tk <= t[0])
            v = y[0];
        else
            v = y[j] + (tk - t[j]) / (t[j + 1] - t[j]) * (y[j + 1] - y[j]);
        out[k] = v;
    }
    return m;
}

void integrate(const double* t, const double* y, size_t n, double y0, double* out) {
#pragma clang fp contract(off)
    double acc = y0;
    if (n == 0) return;
    out[0] = acc;
    for (size_t i = 1; i < n; ++i) {
        acc += 0.5 * (y[i] + y[i - 1]) * (t[i] - t[i - 1]);
        out[i] = acc;
    }
}

double pulse_velocity_change(const double* accel, size_t n, double dt, double cfc) {
#pragma clang fp contract(off)
    if (n < 2) return 0.0;
    std::vector<double> a(accel, accel + n);
    if (cfc > 0.0) a = cfc_filter(a, dt, cfc);
    // trapezoid over the uniform grid k dt (Pulse::velocity, src/signal.rs:190-193)
    double acc = 0.0;
    for (size_t i = 1; i < n; ++i) acc += 0.5 * (a[i] + a[i - 1]) * ((double)i * dt - (double)(i - 1) * dt);
    return acc;
}

void pulse_results_host(const Prepared& P, const double* h, int samples, double t_end, crush_results* out) {
    // the resampling grid of Pulse::from_history ends at floor(t_end / dt)
    // (src/signal.rs:44); the device streams every k dt <= t, which can add
    // the point at t_end when k dt rounds onto it
    if (P.n_pulse > 0 && P.settings.pulse_dt > 0.0)
        samples = std::min(samples, (int)std::floor((t_end - 0.0) / P.settings.pulse_dt) + 1);
    if (P.n_pulse <= 0 || samples <= 0 || !h) {
        out->pulse_samples = 0;
        if (out->groups)
            for (int g = 0; g < P.n_groups; ++g) out->groups[g].pulse_delta_v = 0.0;
        return;
    }
    const int ng = std::min(P.n_groups, 8);
    const int ns = std::min(samples, P.n_pulse);
    out->pulse_samples = ns;
    if (out->groups)
        for (int g = 0; g < P.n_groups; ++g) {
            double dv = 0.0;
            if (g < ng) {
                // HeadOn::run / run_barrier_pulse (src/headon.rs:70-78,
                // src/vehicle.rs:1108-1115): CFC 60, then speed at the start
                // minus speed at the end along the direction of travel
                const double dvx = pulse_velocity_change(h + (size_t)g * P.n_pulse, (size_t)ns, P.settings.pulse_dt,
                                                         60.0);
                dv = P.g_heading[g] < 0 ? dvx : -dvx;
            }
            out->groups[g].pulse_delta_v = dv;
        }
    if (out->pulse && out->pulse_capacity > 0)
        for (int g = 0; g < ng; ++g)
            std::copy(h + (size_t)g * P.n_pulse, h + (size_t)g * P.n_pulse + std::min(ns, out->pulse_capacity),
                      out->pulse + (size_t)g * out->pulse_capacity);
}

std::vector<double> pulse_download(const Prepared& P, const double* dev_pulse) {
    std::vector<double> h((size_t)std::min(P.n_groups, 8) * (size_t)std::max(P.n_pulse, 0));
    if (!h.empty()) HIP_OK(hipMemcpy(h.data(), dev_pulse, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    return h;
}

void pulse_results(const Prepared& P, const double* dev_pulse, int samples, double t_end, crush_results* out) {
    const std::vector<double> h = pulse_download(P, dev_pulse);
    pulse_results_host(P, h.empty() ? nullptr : h.data(), samples, t_end, out);
}

}  // namespace crush

using namespace crush;

extern "C" {

int32_t crush_cfc_filter(const double* x, int32_t n, double dt, double cfc, double* out) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!x || !out))) fail(CRUSH_E_INPUT, "null argument");
        if (n >= 4 && !(dt > 0.0 && cfc > 0.0)) fail(CRUSH_E_INPUT, "dt and cfc must be positive");
        std::vector<double> y = cfc_filter(std::vector<double>(x, x + n), dt, cfc);
        std::copy(y.begin(), y.end(), out);
    });
}

int32_t crush_resample(const double* t, const double* y, int32_t n, double t0, double dt, double t_end,
                       double* out, int32_t cap, int32_t* n_out) {
    return guarded([&] {
        if (n < 1 || !t || !y || !out || !n_out) fail(CRUSH_E_INPUT, "null or empty argument");
        if (!(dt > 0.0) || !(t_end >= t0)) fail(CRUSH_E_INPUT, "need dt > 0 and t_end >= t0");
        const double m = std::floor((t_end - t0) / dt) + 1.0;
        if (m > (double)cap) fail(CRUSH_E_INPUT, "resampled signal has %.0f samples, capacity %d", m, cap);
        *n_out = (int32_t)resample(t, y, (size_t)n, t0, dt, t_end, out, (size_t)cap);
    });
}

int32_t crush_integrate(const double* t, const double* y, int32_t n, double y0, double* out) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!t || !y || !out))) fail(CRUSH_E_INPUT, "null argument");
        integrate(t, y, (size_t)n, y0, out);
    });
}

int32_t crush_pulse_metrics_get(const double* time, const double* accel, int32_t n, double mass, double v0,
                                crush_pulse_metrics* out) {
    return guarded([&] {
        if (n < 1 || !time || !accel || !out) fail(CRUSH_E_INPUT, "null or empty argument");
        // PulseMetrics::new: velocity and displacement by trapezoids, the
        // first maximum of the displacement, the end and minimum velocities
        // and the most negative acceleration (src/signal.rs:230-240)
        std::vector<double> v(n), x(n);
        integrate(time, accel, (size_t)n, v0, v.data());
        integrate(time, v.data(), (size_t)n, 0.0, x.data());
        size_t i_max = 0;
        double max_crush = 0.0;
        for (size_t i = 0; i < (size_t)n; ++i)
            if (x[i] > max_crush) {
                max_crush = x[i];
                i_max = i;
            }
        double v_min = INFINITY, peak = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            v_min = std::fmin(v_min, v[i]);
            peak = std::fmin(peak, accel[i]);
        }
        (void)mass;
        out->t_max_crush = time[i_max];
        out->max_crush = max_crush;
        out->v_end = v[n - 1];
        out->restitution = std::fmax(-v_min / v0, 0.0);
        out->peak_accel = peak;
    });
}

}  // extern "C"
