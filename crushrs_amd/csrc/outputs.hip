// Histories and frames of a solve: host layout of the device history
// records and the one-model path's frames (src/solver.rs:8-62, 114-187,
// 265-357).
#include "outputs.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "hip_util.hpp"

namespace crush {

std::vector<double> history_download(const Prepared& P, const double* d_hist, int n, hipStream_t s) {
    const size_t rl = hist_record_len(P.n_parts, P.n_groups);
    const int kept = std::min(std::max(n, 0), P.n_hist);
    std::vector<double> h(rl * (size_t)kept);
    if (kept > 0) {
        HIP_OK(hipMemcpyAsync(h.data(), d_hist, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    return h;
}

// Frame axes from positions (the host form of phases.hpp frame_axes).
static void frame_axes_host(const double* xo, const double* xx, const double* xp, double* ax) {
    double e1[3], p[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) e1[d] = xx[d] - xo[d];
    if (xp)
        for (int d = 0; d < 3; ++d) p[d] = xp[d] - xo[d];
    double n1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    if (n1 < 1e-300) {
        e1[0] = 1.0;
        e1[1] = e1[2] = 0.0;
        n1 = 1.0;
    }
    for (int d = 0; d < 3; ++d) e1[d] /= n1;
    const double pn = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const double c[3] = {p[1] * e1[2] - p[2] * e1[1], p[2] * e1[0] - p[0] * e1[2], p[0] * e1[1] - p[1] * e1[0]};
    if (std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) < 1e-12 * std::max(pn, 1e-300)) {
        const bool xa = std::fabs(e1[0]) < 0.9;
        p[0] = xa ? 1.0 : 0.0;
        p[1] = xa ? 0.0 : 1.0;
        p[2] = 0.0;
    }
    double e3[3] = {e1[1] * p[2] - e1[2] * p[1], e1[2] * p[0] - e1[0] * p[2], e1[0] * p[1] - e1[1] * p[0]};
    const double n3 = std::sqrt(e3[0] * e3[0] + e3[1] * e3[1] + e3[2] * e3[2]);
    for (int d = 0; d < 3; ++d) e3[d] /= n3;
    const double e2[3] = {e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2], e3[0] * e1[1] - e3[1] * e1[0]};
    for (int d = 0; d < 3; ++d) {
        ax[d] = e1[d];
        ax[3 + d] = e2[d];
        ax[6 + d] = e3[d];
    }
}

// The initial sample (step 0, t = 0): displacement 0, the initial
// velocities, acceleration 0 (every model starts with its contact pairs
// apart and no internal force).
static void initial_record(const Prepared& P, double* rec) {
    rec[0] = 0.0;
    rec[1] = 0.0;
    for (int p = 0; p < P.n_parts; ++p) {
        double m = 0.0, mv[3] = {0, 0, 0}, ke = 0.0;
        for (int i = P.p_nodes_ptr[p]; i < P.p_nodes_ptr[p + 1]; ++i) {
            const int n = P.p_nodes[i];
            m += P.mass[n];
            for (int d = 0; d < 3; ++d) {
                const double vd = P.v0[3 * (size_t)n + d];
                mv[d] += P.mass[n] * vd;
                ke += 0.5 * P.mass[n] * vd * vd;
            }
        }
        double* r = rec + 2 + (size_t)CRUSH_PART_FIELDS * p;
        const double inv = m > 0.0 ? 1.0 / m : 0.0;
        for (int f = 0; f < CRUSH_PART_FIELDS; ++f) r[f] = 0.0;
        r[0] = m;
        for (int d = 0; d < 3; ++d) r[4 + d] = mv[d] * inv;
        r[10] = ke;
    }
    const int na = std::min(P.n_groups, 8);
    for (int g = 0; g < na; ++g) {
        double* r = rec + 2 + (size_t)CRUSH_PART_FIELDS * P.n_parts + (size_t)CRUSH_ACCEL_FIELDS * g;
        for (int f = 0; f < CRUSH_ACCEL_FIELDS; ++f) r[f] = 0.0;
        const int n = P.g_accel[g];
        if (n < 0) {
            r[15] = r[19] = r[23] = 1.0;
            continue;
        }
        for (int d = 0; d < 3; ++d) {
            r[d] = P.X[3 * (size_t)n + d];
            r[6 + d] = P.v0[3 * (size_t)n + d];
        }
        const int o = P.g_frame[3 * g], x = P.g_frame[3 * g + 1], pl = P.g_frame[3 * g + 2];
        if (o >= 0) {
            for (int d = 0; d < 3; ++d) r[12 + d] = P.X[3 * (size_t)o + d];
            frame_axes_host(&P.X[3 * (size_t)o], &P.X[3 * (size_t)x], pl >= 0 ? &P.X[3 * (size_t)pl] : nullptr, r + 15);
        } else {
            for (int d = 0; d < 3; ++d) r[12 + d] = r[d];
            r[15] = r[19] = r[23] = 1.0;
        }
    }
}

void history_results(const Prepared& P, const double* recs, int n_taken, crush_results* out) {
    out->history_samples = 0;
    if (P.settings.history_steps <= 0) return;
    const int kept = std::min(std::max(n_taken, 0), P.n_hist);
    const int total = 1 + kept;
    out->history_samples = total;
    const int cap = std::max(out->history_capacity, 0);
    const size_t rl = hist_record_len(P.n_parts, P.n_groups);
    const int na = std::min(P.n_groups, 8);
    std::vector<double> first(rl);
    initial_record(P, first.data());
    for (int k = 0; k < std::min(total, cap); ++k) {
        const double* r = k == 0 ? first.data() : recs + (size_t)(k - 1) * rl;
        if (out->history_time) out->history_time[k] = r[0];
        if (out->history_step) out->history_step[k] = (int64_t)r[1];
        if (out->history)
            std::memcpy(out->history + (size_t)k * P.n_parts * CRUSH_PART_FIELDS, r + 2,
                        sizeof(double) * CRUSH_PART_FIELDS * P.n_parts);
        if (out->accel_history && na > 0)
            std::memcpy(out->accel_history + (size_t)k * na * CRUSH_ACCEL_FIELDS,
                        r + 2 + (size_t)CRUSH_PART_FIELDS * P.n_parts, sizeof(double) * CRUSH_ACCEL_FIELDS * na);
    }
}

// ------------------------------------------------------------- frames

__global__ __launch_bounds__(256) void k_frame(InstDev D, const int32_t* __restrict__ slot_elem, long long step,
                                               float* __restrict__ fu, float* __restrict__ fpl,
                                               float* __restrict__ fsig, uint8_t* __restrict__ fer,
                                               double* __restrict__ ft) {
    if (D.ctl->step != step) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *ft = D.ctl->t;
    if (i < 3 * D.N) fu[i] = (float)D.u[i];
    if (i < D.E) {
        const int e = slot_elem[i];
        float sig[6], pl;
        element_output(D, i, sig, &pl);
        fpl[e] = pl;
        for (int k = 0; k < 6; ++k) fsig[6 * (size_t)e + k] = sig[k];
        fer[e] = D.eroded[i];
    }
}

FrameStore::~FrameStore() {
    (void)hipFree(u);
    (void)hipFree(plastic);
    (void)hipFree(stress);
    (void)hipFree(eroded);
    (void)hipFree(slot_elem);
    (void)hipFree(t_dev);
}

void FrameStore::init(const Prepared& P, int capacity, hipStream_t s) {
    cap = std::max(capacity, 0);
    if (cap == 0) return;
    const size_t N3 = 3 * (size_t)P.n_nodes, E = (size_t)P.n_elems;
    HIP_OK(hipMalloc(&u, sizeof(float) * N3 * cap));
    HIP_OK(hipMalloc(&plastic, sizeof(float) * E * cap));
    HIP_OK(hipMalloc(&stress, sizeof(float) * 6 * E * cap));
    HIP_OK(hipMalloc(&eroded, E * cap));
    HIP_OK(hipMalloc(&t_dev, sizeof(double) * cap));
    HIP_OK(hipMalloc(&slot_elem, sizeof(int32_t) * E));
    HIP_OK(hipMemcpyAsync(slot_elem, P.slot_elem.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice, s));
    // frame 0: the initial state (no displacement, stress or compaction)
    HIP_OK(hipMemsetAsync(u, 0, sizeof(float) * N3, s));
    HIP_OK(hipMemsetAsync(plastic, 0, sizeof(float) * E, s));
    HIP_OK(hipMemsetAsync(stress, 0, sizeof(float) * 6 * E, s));
    HIP_OK(hipMemsetAsync(eroded, 0, E, s));
    HIP_OK(hipMemsetAsync(t_dev, 0, sizeof(double), s));
}

void FrameStore::record(const InstDev& I, const Prepared& P, int k, long long step, hipStream_t s) {
    if (k >= cap) return;
    const size_t N3 = 3 * (size_t)P.n_nodes, E = (size_t)P.n_elems;
    const long long n = std::max<long long>((long long)N3, (long long)E);
    hipLaunchKernelGGL(k_frame, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, I, slot_elem, step,
                       u + N3 * k, plastic + E * k, stress + 6 * E * k, eroded + E * k, t_dev + k);
    HIP_OK(hipGetLastError());
}

void FrameStore::fill(const Prepared& P, int n_frames, crush_results* out, hipStream_t s) {
    out->frame_samples = n_frames;
    const int m = std::min(std::min(n_frames, cap), std::max(out->frame_capacity, 0));
    if (m <= 0) return;
    const size_t N3 = 3 * (size_t)P.n_nodes, E = (size_t)P.n_elems;
    if (out->frame_time)
        HIP_OK(hipMemcpyAsync(out->frame_time, t_dev, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    if (out->frame_displacement)
        HIP_OK(hipMemcpyAsync(out->frame_displacement, u, sizeof(float) * N3 * m, hipMemcpyDeviceToHost, s));
    if (out->frame_plastic_strain)
        HIP_OK(hipMemcpyAsync(out->frame_plastic_strain, plastic, sizeof(float) * E * m, hipMemcpyDeviceToHost, s));
    if (out->frame_stress)
        HIP_OK(hipMemcpyAsync(out->frame_stress, stress, sizeof(float) * 6 * E * m, hipMemcpyDeviceToHost, s));
    if (out->frame_eroded) HIP_OK(hipMemcpyAsync(out->frame_eroded, eroded, E * m, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
}

}  // namespace crush
